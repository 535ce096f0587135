function [R_jL_mat, Y_mat, flag, relres] = vh_batched_sweep(dx, z_real, z_imag, idxS, st_sparse_precon2, Ae, rhs_vect, freq_all, method)
% All (port, frequency) systems of VoxHenry_executer.m:215-328 in one device
% solve through the gateway's batched tier, for a normal conductor (the
% impedance does not depend on f; two-fluid conductors go through the
% per-frequency loop).  The arguments are the executer's own variables at
% the start of its frequency loop (Ae with the port rows removed, rhs_vect
% truncated, :181-192); method is 'fp64' (pgmres semantics), 'fp64_true' (pgmres, then FP64 FGMRES to
% the true residual for the columns it leaves above tol) or 'mixed'.
% Tolerances follow the executer's rule (:219-223); Y, Z and R + jL are
% assembled as at :311-328.
mu = 4*pi*1e-7;
num_ports = size(rhs_vect, 2);
num_freq = numel(freq_all);
% geometry, Green's tensor and incidence (registered once), impedance
vh_mex('prepare', dx, z_real, z_imag, idxS, (1j*2*pi*freq_all(1)*mu) * st_sparse_precon2, Ae);
jwmu = 1j*2*pi*mu*reshape(freq_all, 1, []);
B = repmat(rhs_vect, 1, num_freq);
pidx = kron(1:num_freq, ones(1, num_ports));
tol = 1e-8 * ones(1, numel(pidx));
tol(freq_all(pidx) < 1e6) = 1e-12;
[X, flag, relres] = vh_mex('solve_batch', B, jwmu, pidx, method, tol);
Y_mat = zeros(num_ports, num_ports, num_freq);
R_jL_mat = zeros(num_ports, num_ports, num_freq);
for f = 1:num_freq
    Xf = X(:, (f-1)*num_ports + (1:num_ports));
    Y_mat(:, :, f) = rhs_vect.' * Xf;                 % sum(rhs_vect(:,kk).*x), unconjugated
    Z = inv(Y_mat(:, :, f));
    R_jL_mat(:, :, f) = abs(real(Z)) + 1j*abs(imag(Z))/(2*pi*freq_all(f));
end
end
