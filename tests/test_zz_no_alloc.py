"""md_lj_forces_noalloc against md_lj_forces on a context holding the same
atoms in the same order: the same kernels over the same staged positions, so
forces, E and W agree bit for bit. Two calls in a row share the library's
fixed scratch array; the second is smaller, so any state the first left
behind (cell counts, partial sums) would show in the second's results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _context_forces(fr, source, lfs):
    from moleculardynamics_amd import native

    ctx = native.Context()
    try:
        ctx.upload(fr.L, np.stack([fr.x, fr.y, fr.z], axis=1))
        ctx.set_lj(1.0, 1.0, 2.5, lfs)
        e, w = ctx.forces(source, native.FP64)
        _, _, f = ctx.download()
    finally:
        ctx.close()
    return f, e, w


def _noalloc_forces(fr, source, lfs, stream=None, forces=True):
    import torch
    from moleculardynamics_amd import native

    n = fr.x.size
    xyz = [torch.tensor(v, dtype=torch.float64, device="cuda:0") for v in (fr.x, fr.y, fr.z)]
    # poisoned outputs: every element must be written
    f = [torch.full((n,), float("nan"), dtype=torch.float64, device="cuda:0") for _ in range(3)]
    torch.cuda.synchronize()
    fp = [t.data_ptr() for t in f] if forces else [0, 0, 0]
    e, w = native.lj_forces_noalloc(source, 1.0, 1.0, 2.5, lfs, n, fr.L,
                                    *[t.data_ptr() for t in xyz], *fp, stream=stream)
    torch.cuda.synchronize()
    return np.stack([t.cpu().numpy() for t in f], axis=1), e, w


@pytest.mark.parametrize("lfs", [False, True], ids=["truncated", "lfs"])
def test_noalloc_cells_matches_context_twice(gpu_ok, orc, lfs):
    from moleculardynamics_amd import native

    for cells in (10, 6):  # 4000 atoms, then 864 (a smaller grid and atom count)
        fr = orc.fcc_lattice(cells, 0.8442, 0.722, 20260000 + cells)
        print(f"noalloc cells: fcc {cells}, n={fr.x.size}, lfs={lfs}", flush=True)
        f_ref, e_ref, w_ref = _context_forces(fr, native.CELLS, lfs)
        f, e, w = _noalloc_forces(fr, native.CELLS, lfs)
        assert np.array_equal(f, f_ref)
        assert e == e_ref and w == w_ref


def test_noalloc_brute_and_energy_only(gpu_ok, orc):
    import torch
    from moleculardynamics_amd import native

    fr = orc.fcc_lattice(5, 0.8442, 0.722, 7)
    f_ref, e_ref, w_ref = _context_forces(fr, native.BRUTE, False)
    s = torch.cuda.Stream()
    f, e, w = _noalloc_forces(fr, native.BRUTE, False, stream=s.cuda_stream)
    assert np.array_equal(f, f_ref)
    assert e == e_ref and w == w_ref
    # energy and virial only, smaller system, after the call above
    fr2 = orc.fcc_lattice(3, 0.8442, 0.722, 8)
    _, e2_ref, w2_ref = _context_forces(fr2, native.BRUTE, False)
    f2, e2, w2 = _noalloc_forces(fr2, native.BRUTE, False, forces=False)
    assert np.all(np.isnan(f2))
    assert e2 == e2_ref and w2 == w2_ref


def test_noalloc_refusals(gpu_ok):
    from moleculardynamics_amd import native

    L = np.array([4.0, 4.0, 4.0])
    with pytest.raises(native.MDError, match="scratch maximum"):
        native.lj_forces_noalloc(native.CELLS, 1.0, 1.0, 2.5, False,
                                 native.NOALLOC_MAX_ATOMS + 1, L, 1, 1, 1)
    with pytest.raises(native.MDError, match="not usable"):
        native.lj_forces_noalloc(native.CELLS, 1.0, 1.0, 2.5, False, 10, L, 1, 1, 1)
