"""FP64 CPU oracle for VoxHenry's solve path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``voxhenry_amd``) imports,
links or executes anything under ``oracle/``; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker / the timed CPU baseline.

It is a plain numpy/scipy restatement of the reference MATLAB code, written from
the reference text (every function cites the file:line it follows):

* ``vhr``       -- ``.vhr`` input reader            (src_pre_process/pre_input_file.m)
* ``topology``  -- index sets, panel IDs, Ae, ports, right-hand sides
                   (pre_define_structure_params.m, lse_generate_panelIDs.m,
                   lse_compute_Ae_matrix.m, lse_assign_exc_grnd_nodes.m,
                   VoxHenry_executer.m:159-193)
* ``green``     -- real-space Green's tensor, circulant embedding, FFT
                   (lse_generate_circulant_tensor.m and src_lin_vie/*)
* ``matvec``    -- lse_matvect_mult.m
* ``precon``    -- lse_sparse_precon_prepare.m / lse_sparse_precon_multiply.m /
                   lse_sparse_schur_compl_multiply.m
* ``gmres``     -- pgmres.m / pgmres_iter.m / fpgmres.m / fpgmres_iter.m
* ``driver``    -- per-frequency loop, Y -> Z -> R, L (VoxHenry_executer.m:215-328)

How each function is pinned (golden data, closed forms, brute force) is listed
in the module docstrings and exercised by ``tests/test_oracle_*.py``.
"""
