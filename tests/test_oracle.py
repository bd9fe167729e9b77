"""The CPU oracle pinned without running the reference (it is Rust; no
toolchain here): the reference's own published numbers and known answers,
closed forms, dense brute force, and the invariants its tests assert
(code/water-core/src/tests_delta3d.rs). Each check would fail for a plausible
mistake (a dropped term, a wrong sign, a transposed face)."""
import struct

import numpy as np
import pytest
import scipy.fft

from fluidisim_amd import workloads as W
from oracle import pyoracle as po


def noise(n, seed):
    """The reference's test noise, tests_delta3d.rs:66-74 (a fixture definition)."""
    s = (seed * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF | 1
    out = np.empty(n, np.float32)
    for q in range(n):
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        out[q] = np.float32(np.float32(s >> 40) / np.float32(16777216.0)) * np.float32(2.0) - np.float32(1.0)
    return out


def zeros_faces(prob):
    return [np.zeros(n, np.float32) for n in prob.face_counts]


# ---------------------------------------------------------------- published numbers
@pytest.mark.parametrize("nx,mg_its,jacobi_its", [(32, 9, 102), (64, 10, 191), (128, 11, 365)])
def test_s385_first_step_iterations_are_the_published_ones(nx, mg_its, jacobi_its):
    """docs/validation/MULTIGRILLE-3D-S385.md:57: first mobile step from rest,
    flat bottom, 9/10/11 iterations with the multigrid V-cycle and
    102/191/365 with Jacobi (examples/delta3d_multigrille.rs:58-103)."""
    prob, scale, k1 = po.s385_problem(nx)
    us, vs, ws = zeros_faces(prob)
    rep, *_ = po.project(prob, us, vs, ws, scale, k1, 20000, po.options(multigrid=1))
    assert rep["iterations"] == mg_its and not rep["degraded"]
    if nx <= 64:
        rep, *_ = po.project(prob, us, vs, ws, scale, k1, 20000, po.options(multigrid=0))
        assert rep["iterations"] == jacobi_its and not rep["degraded"]


@pytest.mark.parametrize("nx,mg_its,jacobi_its", [(32, 9, 117), (64, 10, 219), (128, 11, 444)])
def test_s385_bump_first_step_iterations_are_the_published_ones(nx, mg_its, jacobi_its):
    """MULTIGRILLE-3D-S385.md:59-60: the same first step over the S324 bump
    (0.4 + 0.6 exp(-(((x-3)/1.2)^2 + ((y-1.3)/0.9)^2)) m, examples/delta3d_multigrille.rs:43-47)
    cut into the grid (configure_with_bottom -> cut, delta3d_cut.rs:172-258):
    9/10/11 iterations with multigrid, 117/219/444 with Jacobi. Pins the
    restated cut geometry and the aperture-weighted operator together."""
    prob, scale, k1 = po.s385_problem(nx, bump=True)
    assert prob.solid.any() and ((prob.open_u > 0) & (prob.open_u < 1)).any()
    us, vs, ws = zeros_faces(prob)
    rep, *_ = po.project(prob, us, vs, ws, scale, k1, 20000, po.options(multigrid=1))
    assert rep["iterations"] == mg_its and not rep["degraded"]
    rep, *_ = po.project(prob, us, vs, ws, scale, k1, 20000, po.options(multigrid=0))
    assert rep["iterations"] == jacobi_its and not rep["degraded"]


def test_cut_of_a_flat_bottom_is_the_closed_form():
    """A flat bottom at h (delta3d_cut.rs:172-258 with the 2D formulas of
    delta_projection.rs:795-827): frac and the side apertures are
    clamp((top - h)/dx) per layer, a horizontal face is open iff h < z, the
    walls and the domain's bottom are closed, and cells with no open face
    are solid."""
    nx, ny, nz, dx = 6, 5, 7, 0.5
    h = np.float32(1.3)
    g = po.cut_geometry(nx, ny, nz, dx, np.full(nx * ny, h, np.float32))
    f32 = np.float32
    for k in range(nz):
        top = f32((k + 1) * dx)
        want = np.clip((top - h) / f32(dx), f32(0), f32(1))
        assert np.all(g["frac"].reshape(nz, ny, nx)[k] == want)
        ou = g["open_u"].reshape(nz, ny, nx + 1)[k]
        assert np.all(ou[:, 1:nx] == want) and np.all(ou[:, [0, nx]] == 0)
        ov = g["open_v"].reshape(nz, ny + 1, nx)[k]
        assert np.all(ov[1:ny, :] == want) and np.all(ov[[0, ny], :] == 0)
    ow = g["open_w"].reshape(nz + 1, ny, nx)
    assert np.all(ow[0] == 0)
    for k in range(1, nz + 1):
        assert np.all(ow[k] == (1.0 if h < f32(k * dx) else 0.0))
    assert np.all(g["floor"] == h)


@pytest.mark.parametrize("slope_x,slope_y", [(0.37, 0.0), (0.0, 0.41), (0.23, -0.31)])
def test_cut_volume_is_the_exact_integral(slope_x, slope_y):
    """The fluid volume of each inner column, sum_k frac * dx, is the domain
    height minus the mean bottom over the footprint: for the corner model of
    a linear bottom that mean is the mean of the four corners (the four
    triangles around the centre integrate exactly), through the 2D formula
    (no y dependence) and through the triangles (with one)."""
    nx, ny, nz, dx = 8, 7, 9, 0.25
    f32 = np.float32
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dx
    bottom = (0.6 + slope_x * x[None, :] + slope_y * y[:, None]).astype(f32).ravel()
    g = po.cut_geometry(nx, ny, nz, dx, bottom)
    frac = g["frac"].reshape(nz, ny, nx).astype(np.float64)
    col = frac.sum(axis=0) * dx
    b2 = bottom.reshape(ny, nx).astype(np.float64)
    for j in range(1, ny - 1):
        for i in range(1, nx - 1):
            # corners: mean of the four surrounding column values (interior)
            cs = [0.25 * (b2[jj, ii] + b2[jj, ii - 1] + b2[jj - 1, ii] + b2[jj - 1, ii - 1])
                  for jj, ii in ((j, i), (j, i + 1), (j + 1, i + 1), (j + 1, i))]
            want = nz * dx - np.mean(cs)
            assert abs(col[j, i] - want) <= 1e-5, (i, j, col[j, i], want)


def test_s385_plain_divergence_within_published_bound():
    """MULTIGRILLE-3D-S385.md:76: plain-row divergence <= 5.9e-6 with multigrid."""
    for nx in (32, 64):
        prob, scale, k1 = po.s385_problem(nx)
        us, vs, ws = zeros_faces(prob)
        rep, *_ = po.project(prob, us, vs, ws, scale, k1, 20000, po.options(multigrid=1))
        assert rep["divergence_plain"] <= 5.9e-6


def test_level_count_known_answers():
    """tests_delta3d.rs:1814-1817."""
    assert po.level_count3(32, 16, 24) == 3
    assert po.level_count3(7, 16, 24) == 0
    assert po.level_count3(120, 112, 28) == 2


@pytest.mark.parametrize("dims", [(32, 16, 24), (16, 8, 12), (7, 5, 3)])
def test_hierarchy_float_count(dims):
    """tests_delta3d.rs:1818-1825: z and t fine, then per level 5 cell arrays and the faces."""
    nx, ny, nz = dims
    total = 2 * nx * ny * nz
    x, y, z = nx, ny, nz
    while x % 2 == 0 and y % 2 == 0 and z % 2 == 0 and min(x, y, z) >= 4:
        x, y, z = x // 2, y // 2, z // 2
        total += 5 * x * y * z + (x + 1) * y * z + x * (y + 1) * z + x * y * (z + 1)
    assert po.hierarchy_floats3(*dims) == total


def test_fingerprint_is_fnv1a_64():
    """delta3d.rs:1360-1369: FNV-1a 64 over the little-endian bytes of each
    f32. Published FNV-1a 64 vectors pin the offset basis and prime."""
    assert po.fnv_bytes(b"") == 0xCBF29CE484222325
    assert po.fnv_bytes(b"a") == 0xAF63DC4C8601EC8C
    assert po.fnv_bytes(b"foobar") == 0x85944171F73967E8
    vals = np.array([1.5, -0.0, 3.25e-7], np.float32)
    raw = b"".join(struct.pack("<f", float(v)) for v in vals)
    assert po.fingerprint(vals) == po.fnv_bytes(raw)
    assert po.fingerprint(np.zeros(0, np.float32)) == 0xCBF29CE484222325


def test_roundoff_certificate():
    """delta3d.rs:65-72: gamma_10 = 10u/(1-10u), u = 2^-24."""
    u = 2.0 ** -24
    assert abs(po.lib().orc_roundoff_3d() - 10 * u / (1 - 10 * u)) < 1e-12


# ---------------------------------------------------------------- geometries
def wavy(nx, ny, nz, dx=0.25, seed=1, solids=True, homogeneous=True):
    """A free surface between 2 and nz-1 cells, optional random solid cells."""
    rng = np.random.default_rng(seed)
    cols = nx * ny
    eta = ((2.0 + (nz - 3.5) * rng.random(cols)) * dx).astype(np.float32)
    solid = None
    if solids:
        solid = (rng.random(nx * ny * nz) < 0.08).astype(np.uint8)
    return po.Problem(nx, ny, nz, dx, eta, rho=1025.0, g_eff=9.81, rest=float(nz * dx * 0.5),
                      homogeneous_ghost=homogeneous, solid=solid)


def dense(prob, precision=64):
    n = prob.cells
    M = np.zeros((n, n))
    e = np.zeros(n)
    for c in range(n):
        e[c] = 1.0
        M[:, c] = po.apply64(prob, e) if precision == 64 else po.apply(prob, e.astype(np.float32))
        e[c] = 0.0
    return M


@pytest.mark.parametrize("seed,solids", [(1, False), (2, True), (3, True)])
def test_operator_matrix_is_bitwise_symmetric_and_prec_is_its_inverse_diagonal(seed, solids):
    """tests_delta3d.rs:507-519 (and :352-365): the f32 operator's matrix is
    bitwise symmetric and prec equals 1/diag bitwise; it is positive definite."""
    prob = wavy(4, 3, 7, seed=seed, solids=solids, homogeneous=False)
    M = dense(prob, precision=32).astype(np.float32)
    assert np.array_equal(M.view(np.uint32), M.T.copy().view(np.uint32))
    _, prec = po.assemble(prob, np.zeros(prob.cells, np.float32), 0.0)
    d = np.diag(M)
    inv = np.where(d > 0, np.float32(1.0) / np.where(d > 0, d, 1).astype(np.float32), np.float32(0)).astype(np.float32)
    assert np.array_equal(prec.view(np.uint32), inv.view(np.uint32))
    kind, _ = po.classify(prob)
    wet = kind == 1
    Mw = M[np.ix_(wet, wet)].astype(np.float64)
    assert np.linalg.eigvalsh(Mw).min() > 0


@pytest.mark.parametrize("seed", [4, 5])
def test_projection_matches_a_dense_solve(seed):
    """Brute force: assemble A and b densely from the oracle's own row and
    right-hand side, solve with LAPACK, compare the f64 projection's p; the
    corrected field is then divergence free on plain rows."""
    prob = wavy(6, 5, 7, seed=seed, solids=True, homogeneous=False)
    nu, nv, nw = prob.face_counts
    us, vs, ws = noise(nu, seed), noise(nv, seed + 10), noise(nw, seed + 20)
    scale = -1025.0 / 0.002
    k1 = 0.002 / 1025.0
    kind, ghost = po.classify(prob)
    wet = kind == 1
    div = po.divergence(prob, us, vs, ws)
    b, _ = po.assemble(prob, div, scale)
    M = dense(prob)
    p_exact = np.zeros(prob.cells)
    p_exact[wet] = np.linalg.solve(M[np.ix_(wet, wet)], b[wet].astype(np.float64))
    rep, p32, u, v, w, p64 = po.project(prob, us, vs, ws, scale, k1, 500, po.options(multigrid=1), precision=64)
    assert not rep["degraded"]
    err = np.linalg.norm(p64[wet] - p_exact[wet]) / np.linalg.norm(p_exact[wet])
    assert err < 1e-5
    # the divergence of the corrected field on plain rows
    d_after = po.divergence(prob, u, v, w)
    plain = wet & (ghost == 0)
    assert np.abs(d_after[plain]).max() * prob.dx / max(np.abs(np.concatenate([u, v, w])).max(), 1e-4) <= 1e-5


def test_closed_box_matches_the_dct_closed_form():
    """All-Neumann box: the 7-point operator is diagonalised by DCT-II in x, y
    and z, eigenvalues sum_d (2 - 2 cos(pi m_d / n_d)) / dx^2; the zero mode is
    dropped and mean-free pressures compared."""
    nx, ny, nz, dx = 12, 10, 8, 0.5
    eta = np.full(nx * ny, nz * dx, np.float32)
    prob = po.Problem(nx, ny, nz, dx, eta, homogeneous_ghost=True, closed_lid=True)
    nu, nv, nw = prob.face_counts
    us, vs, ws = noise(nu, 31), noise(nv, 32), noise(nw, 33)
    # walls closed (as set_velocity's close_walls does, delta3d.rs:664-690, plus the lid)
    us3 = us.reshape(nz, ny, nx + 1); us3[:, :, 0] = 0; us3[:, :, -1] = 0
    vs3 = vs.reshape(nz, ny + 1, nx); vs3[:, 0, :] = 0; vs3[:, -1, :] = 0
    ws3 = ws.reshape(nz + 1, ny, nx); ws3[0] = 0; ws3[-1] = 0
    scale, k1 = -1.0, 1.0
    div = po.divergence(prob, us, vs, ws).astype(np.float64).reshape(nz, ny, nx)
    b = scale * div
    b -= b.mean()
    bh = scipy.fft.dctn(b, type=2, norm="ortho")
    lam = np.zeros((nz, ny, nx))
    for axis, n in ((2, nx), (1, ny), (0, nz)):
        m = np.arange(n)
        shape = [1, 1, 1]
        shape[axis] = n
        lam = lam + ((2 - 2 * np.cos(np.pi * m / n)) / dx ** 2).reshape(shape)
    lam[0, 0, 0] = 1.0
    ph = bh / lam
    ph[0, 0, 0] = 0.0
    p_exact = scipy.fft.idctn(ph, type=2, norm="ortho").ravel()
    rep, p32, *_rest, p64 = po.project(prob, us, vs, ws, scale, k1, 500, po.options(multigrid=1), precision=64)
    assert not rep["degraded"]
    p = p64 - p64.mean()
    assert np.linalg.norm(p - p_exact) / np.linalg.norm(p_exact) < 1e-5


def test_flat_surface_at_the_lid_matches_the_dct4_closed_form():
    """Columns full to the lid with an open lid: Neumann at the bottom and a
    Dirichlet zero at the half cell above the top row (the ghost with theta =
    1/2, delta3d_mobile.rs:126-134) make z diagonal in DCT-IV, eigenvalues
    2 - 2 cos(pi (m + 1/2) / nz); x and y stay DCT-II."""
    nx, ny, nz, dx = 10, 8, 9, 0.25
    eta = np.full(nx * ny, nz * dx, np.float32)
    prob = po.Problem(nx, ny, nz, dx, eta, homogeneous_ghost=True, closed_lid=False)
    nu, nv, nw = prob.face_counts
    us, vs, ws = noise(nu, 41), noise(nv, 42), noise(nw, 43)
    us3 = us.reshape(nz, ny, nx + 1); us3[:, :, 0] = 0; us3[:, :, -1] = 0
    vs3 = vs.reshape(nz, ny + 1, nx); vs3[:, 0, :] = 0; vs3[:, -1, :] = 0
    ws3 = ws.reshape(nz + 1, ny, nx); ws3[0] = 0
    scale, k1 = -2.0, 0.5
    b = (scale * po.divergence(prob, us, vs, ws).astype(np.float64)).reshape(nz, ny, nx)
    bh = scipy.fft.dct(scipy.fft.dctn(b, type=2, axes=(1, 2), norm="ortho"), type=4, axis=0, norm="ortho")
    lam = np.zeros((nz, ny, nx))
    lam += ((2 - 2 * np.cos(np.pi * np.arange(nx) / nx)) / dx ** 2).reshape(1, 1, nx)
    lam += ((2 - 2 * np.cos(np.pi * np.arange(ny) / ny)) / dx ** 2).reshape(1, ny, 1)
    lam += ((2 - 2 * np.cos(np.pi * (np.arange(nz) + 0.5) / nz)) / dx ** 2).reshape(nz, 1, 1)
    ph = bh / lam
    p_exact = scipy.fft.idctn(scipy.fft.idct(ph, type=4, axis=0, norm="ortho"), type=2, axes=(1, 2),
                              norm="ortho").ravel()
    rep, p32, *_rest, p64 = po.project(prob, us, vs, ws, scale, k1, 500, po.options(multigrid=1), precision=64)
    assert not rep["degraded"]
    assert np.linalg.norm(p64 - p_exact) / np.linalg.norm(p_exact) < 1e-5


# ---------------------------------------------------------------- invariants
@pytest.mark.parametrize("smoother", [0, 1])
@pytest.mark.parametrize("coarse_op", [0, 1])
def test_vcycle_is_symmetric_and_positive(smoother, coarse_op):
    """tests_delta3d.rs:1840-1872: the V-cycle is symmetric (1e-5 relative)
    and positive, and keeps z = 0 on dry cells, for either smoother and either
    coarse operator."""
    prob = wavy(32, 16, 24, dx=0.25, seed=7, solids=True)
    opts = po.options(multigrid=1, smoother=smoother, coarse_op=coarse_op)
    _, prec = po.assemble(prob, np.zeros(prob.cells, np.float32), 1.0)
    x = noise(prob.cells, 11)
    y = noise(prob.cells, 13)
    x[prec == 0] = 0
    y[prec == 0] = 0
    zx = po.vcycle(prob, x, opts, precision=64)
    zy = po.vcycle(prob, y, opts, precision=64)
    ab, ba = zx @ y, zy @ x
    assert abs(ab - ba) <= 1e-10 * np.sqrt((zx @ zx) * (y @ y))
    assert zx @ x > 0 and zy @ y > 0
    assert np.all(zx[prec == 0] == 0)


def test_vcycle_asymmetry_is_seen_failing():
    """The reference's 'seen failing' control (MULTIGRILLE-3D-S385.md:53):
    one post-sweep instead of two breaks the symmetry and the check notices."""
    prob = wavy(32, 16, 24, dx=0.25, seed=7, solids=False)
    opts = po.options(multigrid=1, smoother=0, post=1)
    _, prec = po.assemble(prob, np.zeros(prob.cells, np.float32), 1.0)
    x, y = noise(prob.cells, 11), noise(prob.cells, 13)
    x[prec == 0] = 0
    y[prec == 0] = 0
    zx = po.vcycle(prob, x, opts, precision=64)
    zy = po.vcycle(prob, y, opts, precision=64)
    # the reference's symmetry tolerance (1e-5 relative) flags it
    assert abs(zx @ y - zy @ x) > 1e-5 * np.sqrt((zx @ zx) * (y @ y))


def test_galerkin_levels_equal_the_reference_ones_in_a_closed_box():
    """With no surface and no solid, P^T A P / 16 is the reference's
    re-discretised operator: same coarse diagonals, same iterations."""
    n = 16
    eta = np.full(n * n, float(n), np.float32)
    prob = po.Problem(n, n, n, 1.0, eta, homogeneous_ghost=True, closed_lid=True)
    for level in range(3):
        k0, d0 = po.coarse_level(prob, level, po.options(coarse_op=0))
        k1, d1 = po.coarse_level(prob, level, po.options(coarse_op=1))
        assert np.array_equal(k0, k1) and np.array_equal(d0, d1)


def test_rest_projects_to_exact_zero():
    """tests_delta3d.rs:180-186."""
    prob = wavy(6, 4, 5, dx=0.5, seed=9, solids=False)
    prob.eta[:] = np.float32(5 * 0.5 * 0.5)
    us, vs, ws = zeros_faces(prob)
    rep, p, u, v, w = po.project(prob, us, vs, ws, -1.0, 1.0, 100, po.options(multigrid=1))
    assert (rep["iterations"], rep["degraded"], rep["residual"], rep["divergence"]) == (0, 0, 0.0, 0.0)
    for a in (p, u, v, w):
        assert not np.any(a.view(np.uint32))


@pytest.mark.parametrize("dims", [(8, 6, 5), (12, 1, 7), (5, 9, 4)])
def test_projection_leaves_a_divergence_free_field(dims):
    """tests_delta3d.rs:151-177 on the mobile operator with the columns full to
    an open lid (the linear-lid operator, delta3d.rs:721-743): divergence
    <= 1e-5, wall faces exactly 0."""
    nx, ny, nz = dims
    dx = 0.25
    eta = np.full(nx * ny, nz * dx, np.float32)
    prob = po.Problem(nx, ny, nz, dx, eta, rho=1025.0, g_eff=9.81, rest=nz * dx, homogeneous_ghost=False)
    nu, nv, nw = prob.face_counts
    u0, v0, w0 = noise(nu, 1), noise(nv, 2), noise(nw, 3)
    u0.reshape(nz, ny, nx + 1)[:, :, [0, -1]] = 0
    v0.reshape(nz, ny + 1, nx)[:, [0, -1], :] = 0
    w0.reshape(nz + 1, ny, nx)[0] = 0
    dt = 0.002
    rep, p, u, v, w = po.project(prob, u0, v0, w0, -1025.0 / dt, dt / 1025.0, 4000, po.options(multigrid=0))
    assert not rep["degraded"] and rep["iterations"] > 0 and rep["divergence"] <= 1e-5
    u3, v3 = u.reshape(nz, ny, nx + 1), v.reshape(nz, ny + 1, nx)
    for face in (u3[:, :, 0], u3[:, :, -1], v3[:, 0, :], v3[:, -1, :]):
        assert not np.any(face.view(np.uint32))


# ---------------------------------------------------------------- 16-bit V-cycle storage
def _storage_samples():
    rng = np.random.default_rng(5)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-9, 6, 20000),
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e6, -1e6, 6.1035156e-05, 6.0e-05, 3e-8, 2.9e-8,
                  5.9604645e-08, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8,
                  3.4e38, -3.4e38, np.inf, -np.inf], np.float32)])
    return x.astype(np.float32)


def test_binary16_rounding_is_numpy_float16():
    """IEEE binary16, nearest-even, subnormals and overflow to inf: numpy's
    float16 conversion is the independent check."""
    x = _storage_samples()
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).astype(np.float32)
    got = po.round_storage(1, x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_bfloat16_rounding_is_torch_bfloat16():
    torch = pytest.importorskip("torch")
    x = _storage_samples()
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = po.round_storage(2, x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_f32_storage_is_the_identity():
    x = _storage_samples()
    assert np.array_equal(po.round_storage(0, x).view(np.uint32), x.view(np.uint32))


@pytest.mark.parametrize("storage", [1, 2])
def test_16bit_storage_vcycle_stays_near_the_f32_one_and_projects_to_tolerance(storage):
    """Rounding the stored V-cycle values perturbs the preconditioner by
    about the storage precision; the projection still reaches 1e-6 in the
    same number of iterations and the same pressure within 1e-5."""
    wl = W.tank(32, 24, 32, seed=9)
    d = wl.domain
    prob = po.Problem(d.nx, d.ny, d.nz, d.dx, wl.eta, rho=wl.rho, g_eff=wl.g_eff, rest=wl.rest,
                      homogeneous_ghost=wl.homogeneous_ghost, closed_lid=wl.closed_lid, solid=wl.solid)
    o = wl.options
    base = dict(multigrid=1, smoother=o["smoother"], max_levels=o["max_levels"], pre=o["pre_sweeps"],
                post=o["post_sweeps"], coarse=o["coarse_sweeps"], coarse_op=o.get("coarse_operator", 0))
    r0, *_, p0 = po.project(prob, wl.us, wl.vs, wl.ws, wl.scale, wl.k1, 500, po.options(**base), precision=64)
    r1, *_, p1 = po.project(prob, wl.us, wl.vs, wl.ws, wl.scale, wl.k1, 500, po.options(storage=storage, **base),
                            precision=64)
    assert r1["iterations"] == r0["iterations"] and r1["residual"] <= 1e-6
    assert np.linalg.norm(p1 - p0) <= 1e-5 * np.linalg.norm(p0)
    _, prec = po.assemble(prob, np.zeros(wl.domain.cells, np.float32), 1.0)
    r = np.random.default_rng(3).standard_normal(wl.domain.cells).astype(np.float32)
    r[prec == 0] = 0
    z0 = po.vcycle(prob, r, po.options(**base))
    z1 = po.vcycle(prob, r, po.options(storage=storage, **base))
    rel = np.abs(z1 - z0).max() / np.abs(z0).max()
    assert 0 < rel < (2e-2 if storage == 2 else 2e-3)
