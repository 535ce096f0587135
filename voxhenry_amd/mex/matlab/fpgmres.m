function [x,flag,relres,iter,resvec] = fpgmres(A,b,restart,tol,maxit,R1,varargin)
% Shadow of src_iterative_solver/fpgmres.m for the executer's schur_gmres
% call fpgmres(fACPU, rhs, inner_it, tol, outer_it, fPCPU): flexible GMRES on
% the device.  Any other use goes to the original, kept on the path under the
% name fpgmres_orig.m.
if nargin == 6 && isa(A, 'function_handle') && isa(R1, 'function_handle') && ...
        ~isempty(strfind(func2str(A), 'fACPU')) && ~isempty(strfind(func2str(R1), 'fPCPU'))
    [x,flag,relres,iter,resvec] = vh_mex('fpgmres', b, restart, tol, maxit);
else
    [x,flag,relres,iter,resvec] = fpgmres_orig(A,b,restart,tol,maxit,R1,varargin{:});
end
end
