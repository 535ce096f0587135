function lse_sparse_precon_prepare(dx,z_real,z_imag,idxS,st_sparse_precon,nodeid_4_grnd,nodeid_4_injectcurr,Ae)
% Shadow of src_lse_formation/lse_sparse_precon_prepare.m (put this folder
% first on the MATLAB path).  Registers the geometry with the GPU library the
% first time it is seen (Green's tensor generated on the device, incidence Ae
% with the port rows already removed by the executer), sets this frequency's
% impedance and prepares the preconditioner on the device.  j*omega*mu is
% recovered as st_sparse_precon(1) / st_sparse_precon2(1).
global fl_precon_type
if isempty(fl_precon_type)
    t = 'schur_invert';
else
    t = fl_precon_type;
end
vh_mex('prepare', dx, z_real, z_imag, idxS, st_sparse_precon, Ae, t);
end
