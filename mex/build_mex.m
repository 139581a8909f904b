% build_mex.m -- builds the four MEX gateways onto libse2p.so (run from the
% repository root in MATLAB, after `make -C fasttools2d_amd/csrc`).  The
% binaries replace PeriodicEwald's bin/mex_stokes_*.mexa64 one for one; the
% drivers StokesSLP_ewald_2p.m / StokesDLP_ewald_2p.m call them unchanged.
lib = fullfile(pwd, 'fasttools2d_amd', 'lib');
names = {'mex_stokes_slp_real', 'mex_stokes_slp_kspace', ...
         'mex_stokes_dlp_real', 'mex_stokes_dlp_kspace'};
for k = 1:numel(names)
    mex('-R2017b', ['-I' fullfile(pwd, 'include')], ['-I' fullfile(pwd, 'mex')], ...
        ['-L' lib], '-lse2p', ['LDFLAGS=$LDFLAGS -Wl,-rpath,' lib], ...
        '-outdir', fullfile(pwd, 'mex'), fullfile('mex', [names{k} '.cpp']));
end
