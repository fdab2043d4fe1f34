"""GPU parity of the LDU kernels against the oracle: SpMV (A and A^T), packed
SpMV, DIC/DILU factors and the Laplacian assembly.  Integer-exact work and the
op-by-op float64 recurrences are compared bit for bit."""

import numpy as np
import pytest
import torch

from oracle import ldu as oldu
from oracle import precond as opre
from pyopenfoam_amd import LduMatrix
from pyopenfoam_amd.solvers import DICPreconditioner, DILUPreconditioner
from tests.helpers import bits_equal, system

pytestmark = pytest.mark.gpu

CASES = [(4, 0.0, 0.0), (8, 0.2, 1.0), (16, 0.2, 1.0), (32, 0.0, 0.0), (32, 0.2, 1.0)]


def _matrix(c, d, lo, up, dev):
    m = LduMatrix(c.n_cells, torch.from_numpy(c.owner), torch.from_numpy(c.neighbour), device=dev)
    m.diag = torch.from_numpy(d)
    m.lower = torch.from_numpy(lo)
    m.upper = torch.from_numpy(up)
    return m


@pytest.mark.parametrize("n,jit,sig", CASES)
def test_spmv_bit_exact(gpu_device, n, jit, sig):
    c, d, lo, up = system(n, jit, sig)
    print(f"spmv n={n} jitter={jit} sigma={sig}", flush=True)
    m = _matrix(c, d, lo, up, gpu_device)
    x = np.random.default_rng(1).standard_normal(c.n_cells)
    xt = torch.from_numpy(x).to(gpu_device)
    y = m.Ax(xt).cpu().numpy()
    yt = m.Ax_T(xt).cpu().numpy()
    yp = m.Ax_packed(xt).cpu().numpy()
    ypt = m.Ax_packed(xt, transpose=True).cpu().numpy()
    assert bits_equal(y, oldu.ax(d, lo, up, c.owner, c.neighbour, x))
    assert bits_equal(yt, oldu.ax_t(d, lo, up, c.owner, c.neighbour, x))
    assert bits_equal(yp, y)
    # packed transpose = Ax of the LDU with lower/upper swapped
    assert bits_equal(ypt, oldu.ax(d, up, lo, c.owner, c.neighbour, x))


def test_spmv_known_answers(gpu_device):
    # reference tests/unit/core/test_ldu_matrix.py:151-202
    own, nei = torch.tensor([0]), torch.tensor([1])
    m = LduMatrix(2, own, nei, device=gpu_device)
    m.diag = torch.tensor([4.0, 4.0])
    m.lower = torch.tensor([-1.0])
    m.upper = torch.tensor([-2.0])
    y = m.Ax(torch.tensor([1.0, 3.0], dtype=torch.float64))
    assert y.cpu().tolist() == [1.0, 10.0]
    m3 = LduMatrix(3, torch.tensor([0, 1]), torch.tensor([1, 2]), device=gpu_device)
    m3.diag = torch.tensor([4.0, 6.0, 4.0])
    m3.lower = torch.tensor([-1.0, -1.0])
    m3.upper = torch.tensor([-1.0, -1.0])
    assert m3.Ax(torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64)).cpu().tolist() == [2.0, 8.0, 10.0]
    z = LduMatrix(2, own, nei, device=gpu_device)
    assert z.Ax(torch.tensor([5.0, 7.0], dtype=torch.float64)).cpu().tolist() == [0.0, 0.0]
    with pytest.raises(ValueError):
        m.Ax(torch.zeros(3, dtype=torch.float64))
    with pytest.raises(TypeError):
        m.Ax(torch.zeros(2, dtype=torch.int64))


@pytest.mark.parametrize("n,jit,sig", CASES)
def test_dic_factor_bit_exact(gpu_device, n, jit, sig):
    c, d, lo, up = system(n, jit, sig)
    print(f"dic n={n}", flush=True)
    m = _matrix(c, d, lo, up, gpu_device)
    rd = DICPreconditioner(m).diagonal.cpu().numpy()
    assert bits_equal(rd, opre.dic_build(d, up, c.owner, c.neighbour))
    # adjoint factor: DIC of A^T uses the swapped coefficients
    rdt = DICPreconditioner(m, transpose=True).diagonal.cpu().numpy()
    assert bits_equal(rdt, opre.dic_build(d, lo, c.owner, c.neighbour))


@pytest.mark.parametrize("n,jit,sig", CASES)
def test_dilu_factor_bit_exact(gpu_device, n, jit, sig):
    c, d, lo, up = system(n, jit, sig)
    print(f"dilu n={n}", flush=True)
    m = _matrix(c, d, lo, up, gpu_device)
    rd = DILUPreconditioner(m).diagonal.cpu().numpy()
    assert bits_equal(rd, opre.dilu_build(d, lo, up, c.owner, c.neighbour))


@pytest.mark.parametrize("lds", ["1", "0"])
def test_dilu_sweep_forms(gpu_device, monkeypatch, lds):
    """The DILU sweep with the diagonal held in LDS (one slot per cell for the
    window of face levels that touch it, slots reused) and on the global
    diagonal (PFM_DILU_LDS=0): both bit-exact against the oracle's sequential
    recurrence, on a 24^3 jittered mesh and on random banded topologies, all
    past the level count below which the factor is launched level by level:
    faces shuffled in order and orientation (a level of 754 faces: the
    1024-thread form of the LDS sweep), sorted with either orientation (the
    512-thread form), and led by 1,400 disjoint faces (a level wider than the
    LDS sweep takes: the factor falls back to the global diagonal)."""
    monkeypatch.setenv("PFM_DILU_LDS", lds)
    print(f"dilu sweep, PFM_DILU_LDS={lds}", flush=True)
    c, d, lo, up = system(24, 0.2, 1.0)
    m = _matrix(c, d, lo, up, gpu_device)
    assert m.addressing.info.n_dilu_levels > 64
    assert bits_equal(DILUPreconditioner(m).diagonal.cpu().numpy(),
                      opre.dilu_build(d, lo, up, c.owner, c.neighbour))
    rng = np.random.default_rng(29)
    n, m = 3000, 20000
    a = rng.integers(0, n, m)
    b = np.clip(a + rng.integers(1, 40, m), 0, n - 1)
    flip = rng.random(m) < 0.5
    own, nei = np.where(flip, b, a), np.where(flip, a, b)
    keep = own != nei
    own, nei = own[keep], nei[keep]
    # shuffled, sorted by cell pair with either orientation, led by disjoint faces
    key = np.minimum(own, nei) * n + np.maximum(own, nei)
    pairs = np.arange(1400)
    wide = (np.concatenate([2 * pairs, own]), np.concatenate([2 * pairs + 1, nei]))
    for o, q in ((own, nei), (own[np.argsort(key, kind="stable")], nei[np.argsort(key, kind="stable")]),
                 wide):
        lo2 = -rng.uniform(0.1, 1.0, o.shape[0])
        up2 = -rng.uniform(0.1, 1.0, o.shape[0])
        d2 = np.bincount(o, -lo2, n) + np.bincount(q, -up2, n) + 0.5
        m2 = LduMatrix(n, torch.from_numpy(o), torch.from_numpy(q), device=gpu_device)
        m2.diag, m2.lower, m2.upper = (torch.from_numpy(d2), torch.from_numpy(lo2),
                                       torch.from_numpy(up2))
        assert m2.addressing.info.n_dilu_levels > 64
        assert bits_equal(DILUPreconditioner(m2).diagonal.cpu().numpy(),
                          opre.dilu_build_py(d2, lo2, up2, o, q))


def test_factor_unordered_faces(gpu_device):
    """Faces not in upper-triangular order (owner > neighbour, shuffled): the
    level schedules must still reproduce the sequential recurrences."""
    rng = np.random.default_rng(3)
    n = 300
    own = rng.integers(0, n, 2000)
    nei = rng.integers(0, n, 2000)
    keep = own != nei
    own, nei = own[keep], nei[keep]
    lo = -rng.uniform(0.1, 1.0, own.shape[0])
    up = -rng.uniform(0.1, 1.0, own.shape[0])
    d = np.bincount(own, -lo, n) + np.bincount(nei, -up, n) + 0.5
    m = LduMatrix(n, torch.from_numpy(own), torch.from_numpy(nei), device=gpu_device)
    m.diag, m.lower, m.upper = torch.from_numpy(d), torch.from_numpy(lo), torch.from_numpy(up)
    assert bits_equal(DICPreconditioner(m).diagonal.cpu().numpy(), opre.dic_build_py(d, up, own, nei))
    assert bits_equal(DILUPreconditioner(m).diagonal.cpu().numpy(),
                      opre.dilu_build_py(d, lo, up, own, nei))
    x = rng.standard_normal(n)
    assert bits_equal(m.Ax(torch.from_numpy(x).to(gpu_device)).cpu().numpy(),
                      oldu.ax(d, lo, up, own, nei, x))
    # rows with more than four couplings per group take the row-slice kernels'
    # one-slot-at-a-time path (packed SpMV and the PCG's SpMV)
    assert m.addressing.info.max_lower > 4 and m.addressing.info.max_upper > 4
    assert bits_equal(m.Ax_packed(torch.from_numpy(x).to(gpu_device)).cpu().numpy(),
                      oldu.ax(d, lo, up, own, nei, x))
    from oracle import krylov as okry
    from pyopenfoam_amd import PCGSolver
    b = rng.uniform(-1.0, 1.0, n)
    s = PCGSolver(tolerance=1e-300, rel_tol=1e-12, max_iter=40, preconditioner="none")
    xs, it, _ = s(m, torch.from_numpy(b), torch.zeros(n, dtype=torch.float64))
    xo, io, _, _ = okry.pcg(lambda v: oldu.ax(d, lo, up, own, nei, v), None, b, np.zeros(n),
                            tolerance=1e-300, rel_tol=1e-12, max_iter=40,
                            dot=okry.dot_device(n))
    assert it == io and bits_equal(xs.cpu().numpy(), xo)


@pytest.mark.parametrize("n,band_lo,band_hi", [(36, 32767, 65535), (42, 65535, None)])
def test_spmv_and_pcg_wide_bandwidth(gpu_device, n, band_lo, band_hi):
    """Randomly permuted meshes without the RCM renumbering (every owner still
    precedes its neighbour).  At 36^3 the column offsets exceed the int16
    range but stay within 65535: the SpMV and PCG kernels read the paired
    offsets as unsigned magnitudes.  At 42^3 they exceed 65535 too and the
    kernels read the int32 columns.  Both stay bit-exact (the renumbered cases
    above read signed-range offsets)."""
    from oracle import krylov as okry
    from pyopenfoam_amd import PCGSolver
    from tests.helpers import make_case
    c = make_case(n, jitter=0.0, sigma=1.0, seed=2, renumber=False)
    band = int(np.abs(c.owner - c.neighbour).max())
    assert band > band_lo and (band_hi is None or band <= band_hi)
    assert (c.neighbour > c.owner).all()
    nfi = c.mesh.n_internal_faces
    d, lo, up = oldu.laplacian(c.gamma, c.geo.mag_sf[:nfi], c.geo.delta[:nfi], c.geo.cell_volumes,
                               c.owner, c.neighbour)
    d = oldu.dirichlet_diag(d, c.dir_gamma, c.dir_mag, c.dir_delta, c.geo.cell_volumes,
                            c.dir_cells)
    m = _matrix(c, d, lo, up, gpu_device)
    x = np.random.default_rng(1).standard_normal(c.n_cells)
    yp = m.Ax_packed(torch.from_numpy(x).to(gpu_device)).cpu().numpy()
    mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
    assert bits_equal(yp, mv(x))
    s = PCGSolver(tolerance=1e-300, rel_tol=1e-6, max_iter=400, preconditioner="none")
    xs, it, _ = s(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    xo, io, _, info = okry.pcg(mv, None, c.b, np.zeros(c.n_cells), tolerance=1e-300, rel_tol=1e-6,
                               max_iter=400, dot=okry.dot_device(c.n_cells))
    assert it == io and bits_equal(xs.cpu().numpy(), xo)


def test_strided_coefficients_match_contiguous(gpu_device):
    """Coefficient tensors that are strided views (the setters alias them, as
    the reference's do) give the same products, factors and solves as
    contiguous copies: each contiguous copy a kernel reads stays alive until
    the kernel has run, so two copies never share one allocator block."""
    from pyopenfoam_amd import GAMGSolver
    c, d, lo, up = system(16, 0.2, 1.0)
    dev = gpu_device
    ref = _matrix(c, d, lo, up, dev)

    def strided(a):
        big = torch.zeros(2 * a.shape[0], dtype=torch.float64, device=dev)
        big[1::2] = torch.from_numpy(a).to(dev)
        return big[1::2]

    m = LduMatrix(c.n_cells, torch.from_numpy(c.owner), torch.from_numpy(c.neighbour), device=dev)
    m.diag, m.lower, m.upper = strided(d), strided(lo), strided(up)
    assert not m.lower.is_contiguous() and not m.upper.is_contiguous()
    x = torch.from_numpy(np.random.default_rng(3).standard_normal(c.n_cells)).to(dev)
    for f in ("Ax", "Ax_T", "Ax_packed"):
        assert bits_equal(getattr(m, f)(x).cpu().numpy(), getattr(ref, f)(x).cpu().numpy()), f
    for P in (DICPreconditioner, DILUPreconditioner):
        assert bits_equal(P(m).diagonal.cpu().numpy(), P(ref).diagonal.cpu().numpy()), P.__name__
    b = torch.from_numpy(c.b).to(dev)
    z = torch.zeros(c.n_cells, dtype=torch.float64, device=dev)
    out = []
    for mat in (m, ref):
        solver = GAMGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=8)
        xs, it, _ = solver(mat, b, z)
        out.append((xs.cpu().numpy(), it))
    assert out[0][1] == out[1][1] and bits_equal(out[0][0], out[1][0])


def test_bandwidth_probe_reads_and_copies(gpu_device):
    """pfm_bandwidth_probe: the read mode's per-wave partials add up to the
    buffer's sum (every element read once), the copy mode copies; sizes that
    are not multiples of 16 bytes are rejected before any launch."""
    from pyopenfoam_amd import _lib
    from pyopenfoam_amd.core.device import current_stream
    L = _lib.lib()
    n = (1 << 20) + 6  # doubles: 16-byte multiple, not a multiple of the grid
    a = torch.arange(n, dtype=torch.float64, device=gpu_device)
    out = torch.zeros(int(L.pfm_bandwidth_probe_out_bytes()) // 8, dtype=torch.float64,
                      device=gpu_device)
    s = current_stream(gpu_device)
    _lib.check(L.pfm_bandwidth_probe(0, _lib.ptr(a), _lib.ptr(out), 8 * n, s), "read")
    assert out.sum().item() == n * (n - 1) / 2  # integers below 2^53: exact
    b = torch.empty_like(a)
    _lib.check(L.pfm_bandwidth_probe(1, _lib.ptr(a), _lib.ptr(b), 8 * n, s), "copy")
    assert torch.equal(a, b)
    assert L.pfm_bandwidth_probe(0, _lib.ptr(a), _lib.ptr(out), 8 * n - 8, s) == _lib.PFM_ERR_ARG


@pytest.mark.parametrize("n,jit,sig", [(8, 0.2, 1.0), (16, 0.2, 1.0)])
def test_batched_and_sparse_products(gpu_device, n, jit, sig):
    """Ax_batched / Ax_sparse (ldu_matrix.py:345-381): every column of A X for
    k right-hand sides bit-identical to Ax of that column (and so to the
    oracle's scatter-order product); a 1-D input is Ax itself.  The sparse
    forms (to_sparse_coo / _csr / _csr_cached, :278-333) hold exactly the
    dense matrix's entries; the CSR cache is reused until a coefficient
    setter invalidates it."""
    c, d, lo, up = system(n, jit, sig)
    m = _matrix(c, d, lo, up, gpu_device)
    X = np.random.default_rng(7).standard_normal((c.n_cells, 5))
    Y = m.Ax_batched(torch.from_numpy(X)).cpu().numpy()
    assert Y.shape == X.shape
    for j in range(5):
        ref = oldu.ax(d, lo, up, c.owner, c.neighbour, X[:, j])
        assert bits_equal(Y[:, j], ref), j
        assert bits_equal(Y[:, j], m.Ax(torch.from_numpy(X[:, j].copy())).cpu().numpy())
    x1 = torch.from_numpy(X[:, 2].copy())
    assert bits_equal(m.Ax_sparse(x1).cpu().numpy(), Y[:, 2])
    with pytest.raises(ValueError):
        m.Ax_batched(torch.zeros(c.n_cells + 1, 2, dtype=torch.float64))
    with pytest.raises(TypeError):
        m.Ax_batched(torch.zeros(c.n_cells, 2, dtype=torch.int64))
    dense = oldu.dense(d, lo, up, c.owner, c.neighbour)
    coo = m.to_sparse_coo()
    assert coo.is_coalesced() and coo.shape == (c.n_cells, c.n_cells)
    assert np.array_equal(coo.to_dense().cpu().numpy(), dense)
    csr = m.to_sparse_csr_cached()
    assert m.to_sparse_csr_cached() is csr
    assert np.array_equal(csr.to_dense().cpu().numpy(), dense)
    m.diag = torch.from_numpy(d * 2.0)
    assert m.to_sparse_csr_cached() is not csr


def test_fv_matrix_relax_sources_and_boundary_hooks(gpu_device):
    """FvMatrix.relax (fv_matrix.py:192-235): diag / alpha and
    source + (1 - alpha) * diag' * field_old, bit for bit as those ops; alpha
    = 1 changes nothing; alpha outside (0, 1] raises ValueError.
    add_explicit_source adds a tensor or a scalar; add_boundary_contribution
    takes the new diagonal and source from bc.matrix_contributions."""
    from pyopenfoam_amd.core.fv_matrix import FvMatrix
    c, d, lo, up = system(8, 0.2, 1.0)
    dev = gpu_device
    m = FvMatrix(c.n_cells, torch.from_numpy(c.owner), torch.from_numpy(c.neighbour), device=dev)
    m.diag, m.lower, m.upper = (torch.from_numpy(v) for v in (d, lo, up))
    m.source = torch.from_numpy(c.b)
    phi = np.random.default_rng(3).standard_normal(c.n_cells)
    m.relax(torch.from_numpy(phi), 1.0)
    assert m.relaxation_factor == 1.0 and bits_equal(m.diag.cpu().numpy(), d)
    a = 0.7
    m.relax(torch.from_numpy(phi), a)
    dn = d / a
    assert m.relaxation_factor == a
    assert bits_equal(m.diag.cpu().numpy(), dn)
    assert bits_equal(m.source.cpu().numpy(), c.b + ((1.0 - a) * dn) * phi)
    for bad in (0.0, -0.5, 1.5):
        with pytest.raises(ValueError):
            m.relax(torch.from_numpy(phi), bad)
    s0 = m.source.cpu().numpy()
    m.add_explicit_source(torch.from_numpy(phi))
    m.add_explicit_source(torch.tensor(2.0, dtype=torch.float64))
    assert bits_equal(m.source.cpu().numpy(), (s0 + phi) + 2.0)

    class PenaltyBC:  # a boundary condition object as the reference's BCs expose it
        def matrix_contributions(self, field, n_cells, diag, source):
            assert field.shape[0] == n_cells and not field.any()
            return diag + 1.0, source - 1.0

    d0, s1 = m.diag.cpu().numpy(), m.source.cpu().numpy()
    m.add_boundary_contribution(PenaltyBC())
    assert bits_equal(m.diag.cpu().numpy(), d0 + 1.0)
    assert bits_equal(m.source.cpu().numpy(), s1 - 1.0)
