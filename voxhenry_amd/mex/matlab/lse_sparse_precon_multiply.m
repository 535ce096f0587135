function [JOut_full_out]=lse_sparse_precon_multiply(JOut_full_in,Ae,nodeid_4_grnd,nodeid_4_injectcurr,prectol)
% Shadow of src_lse_formation/lse_sparse_precon_multiply.m (schur_invert and
% schur_gmres branches), applied on the device.
if nargin < 5
    JOut_full_out = vh_mex('precon_apply', JOut_full_in);
else
    JOut_full_out = vh_mex('precon_apply', JOut_full_in, prectol);
end
end
