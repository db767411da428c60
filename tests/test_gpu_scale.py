"""Full-size parity on one MI355X against size-independent closed forms.

* 2-D Laplacian on a 4096 x 4096 grid (n = 16.8M), t = 10, m = 40, tol 1e-8,
  Expokit step control, column by column, with s-step blocks of 12 after 6
  columns, and with the bench's shape (16 after 8): y against the DST-I
  closed form.
* 3-D convection-diffusion on 256^3 (n = 16.8M, Pe = 50), t = 0.5, m = 50,
  FP32-stored and FP64 values, and with s-step blocks (12 after 6; the
  bench's 14 after 8, FP64 and FP32 values): y against the Kronecker closed
  form.
"""
import numpy as np
import pytest

from closed_forms import kron3_expmv, laplacian_2d_expmv, rel_err
from conftest import gen, mek

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sstep,first", [(0, -1), (12, 6), (-1, -1)])
def test_laplacian_4096_expokit(gpu, sstep, first):
    N, t = 4096, 10.0
    sstep, first = mek.sstep_plan(40, sstep, first)
    print(f"[scale] laplacian 4096^2 build (sstep {sstep}, first {first})", flush=True)
    rp, ci, va = gen.laplacian_2d(N)
    v = gen.normal_vector(N * N, seed=0)
    R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=40, t=t, tol=1e-8, max_restarts=0,
                   step_control="expokit", gs_mode=2 if sstep else -1, sstep=sstep, sstep_first=first)
    del rp, ci, va
    print("[scale] laplacian 4096^2 solve", flush=True)
    y = R.expmv(v)
    s = R.stats()
    print(f"[scale] laplacian steps={s['accepted']} attempts={s['attempts']} ms={s['solve_ms']:.1f}", flush=True)
    err = rel_err(y, laplacian_2d_expmv(N, t, v))
    print(f"[scale] laplacian rel err {err:.3e}", flush=True)
    assert err <= 1e-8
    R.close()


@pytest.mark.parametrize("f32,sstep,first", [(True, 0, -1), (False, 0, -1), (False, 12, 6), (False, -1, -1),
                                             (True, -1, -1)])
def test_convdiff_256(gpu, f32, sstep, first):
    N, t = 256, 0.5
    sstep, first = mek.sstep_plan(50, sstep, first)
    print(f"[scale] convdiff 256^3 f32={f32} build (sstep {sstep}, first {first})", flush=True)
    rp, ci, va = gen.convdiff_3d(N)
    (lx, dx, ux), (ly, dy, uy), (lz, dz, uz) = gen.convdiff_3d_coeffs()
    v = gen.normal_vector(N ** 3, seed=0)
    R = mek.Runner(mek.CSR.from_arrays(rp, ci, va), m=50, t=t, tol=1e-8, max_restarts=0,
                   step_control="expokit", values_f32=f32, gs_mode=2 if sstep else -1, sstep=sstep,
                   sstep_first=first)
    del rp, ci, va
    print("[scale] convdiff solve", flush=True)
    y = R.expmv(v)
    s = R.stats()
    print(f"[scale] convdiff steps={s['accepted']} attempts={s['attempts']} ms={s['solve_ms']:.1f}", flush=True)
    ref = kron3_expmv(gen.tridiag_1d(N, lz, dz, uz), gen.tridiag_1d(N, ly, dy, uy),
                      gen.tridiag_1d(N, lx, dx, ux), t, v)
    err = rel_err(y, ref)
    print(f"[scale] convdiff rel err {err:.3e}", flush=True)
    assert err <= 1e-8
    R.close()
