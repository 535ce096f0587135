function [schur_out]=lse_sparse_schur_compl_multiply(schur_in, Ae, A_inv)
% Shadow of src_lse_formation/lse_sparse_schur_compl_multiply.m:
% Ae*A_inv*Ae'*schur_in with the device copies of Ae and A_inv.
schur_out = vh_mex('schur_multiply', schur_in);
end
