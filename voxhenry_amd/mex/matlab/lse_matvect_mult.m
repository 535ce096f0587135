function [JOut_full]=lse_matvect_mult(JIn0, fN_all, Ae, z_real, z_imag, dx, idx, nodeid_4_grnd, nodeid_4_injectcurr)
% Shadow of src_lse_formation/lse_matvect_mult.m: the operator data were
% registered by the shadow lse_sparse_precon_prepare of this frequency; the
% arguments are kept for the caller's signature only.
JOut_full = vh_mex('matvec', JIn0);
end
