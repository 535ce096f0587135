function [x,flag,relres,iter,resvec] = pgmres(A,b,restart,tol,maxit,L1,L2,R1,R2,x0)
% Shadow of src_iterative_solver/pgmres.m for the executer's call
% pgmres(fACPU, rhs, inner_it, tol, outer_it, fPCPU): the whole solve runs on
% the device with the operator and preconditioner of the last
% lse_sparse_precon_prepare.  Any other use goes to the original, kept on the
% path under the name pgmres_orig.m.
if nargin == 6 && isa(A, 'function_handle') && isa(L1, 'function_handle') && ...
        ~isempty(strfind(func2str(A), 'fACPU')) && ~isempty(strfind(func2str(L1), 'fPCPU'))
    [x,flag,relres,iter,resvec] = vh_mex('pgmres', b, restart, tol, maxit);
else
    args = {A,b,restart,tol,maxit};
    if nargin >= 6, args{end+1} = L1; end
    if nargin >= 7, args{end+1} = L2; end
    if nargin >= 8, args{end+1} = R1; end
    if nargin >= 9, args{end+1} = R2; end
    if nargin >= 10, args{end+1} = x0; end
    [x,flag,relres,iter,resvec] = pgmres_orig(args{:});
end
end
