"""CPU FP64 oracle for the 3D-RISM solute-in-solvent path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``bgy3d_amd``) imports,
links or calls anything under ``oracle/``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
use it, and only as the checker (or the timed CPU baseline), never as the
thing measured or shipped.

The reference (PETSc/FFTW-MPI/Fortran/Guile) cannot be built here, so this
package is a line-by-line restatement of the reference algorithm in numpy:

* ``closures``  -- closures.f90 (closure, closure1, form)
* ``rism1d``    -- rism.f90 rism_vv / rism_uv / rism_solvent_renorm,
                   drism.f90 dipole_correction, fft.f90 mkgrid/fourier_rows
* ``rism3d``    -- hnc3d.c hnc3d_solute_solve / iterate_t1 / jacobian_t1 /
                   solvent_kernel_rism / chempot, bgy3d-solutes.c fields,
                   bgy3d-vec.c interp / vec_ktab / bgy3d_vec_fft_trans
* ``nksolve``   -- a Newton-GMRES solver standing in for PETSc SNES
                   NEWTONLS + KSPGMRES (bgy3d-snes.c:161-346); only the
                   fixed point matters for parity, not the iteration path.

Pinning (what makes this oracle trustworthy) is listed per function in the
module docstrings and exercised in ``tests/test_oracle_*.py``.
"""
