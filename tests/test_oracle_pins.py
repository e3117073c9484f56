"""Independent pins for the oracle's sampled-filament path, graded
discretization and skin-depth counts, and the product's discretization
checked bit for bit against the oracle's on graded grids.

Restated from the reference's own tests:
* tests/inductance.rs:565-612 (general orientation against a fine
  sixfold quadrature, near skew and rotated-parallel pairs at 2e-6),
  :641-700 (touching bends and crossings against a finer sampling of the
  line formula, 1e-3), :706-759 (the parallel limit against the bar closed
  form, 1e-5, tilts down to 2e-6 rad);
* filament.rs:935-1080 (graded grids: tiling, geometric progression,
  symmetry, the hand-computed fixture, rigid motion, invalid ratios) and
  the doc values of graded_surface_extent (filament.rs:243-258);
* solve.rs:388-402 (SkinDepthGrading::count_for) against a brute-force
  count built here from the closed-form geometric sums.

The independent references (numpy Gauss-Legendre sums, closed-form
geometric series, rotation matrices) share no code with the oracle.
"""
import math

import numpy as np
import pytest

import oracle
from fasterhenry_amd import discretize_packed, fixtures, model
from fasterhenry_amd.errors import InvalidGrading

K = 1e-7  # mu0 / 4 pi
COPPER = 5.8e7


@pytest.fixture(autouse=True)
def _portable():
    oracle.set_portable_math(True)
    yield


def rel(a, b):
    return abs((a - b) / b)


def gl01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def single(a, b, w, h, width_dir=None, grid=None):
    g = model.Geometry()
    na = g.add_node(model.Node(*map(float, a)))
    nb = g.add_node(model.Node(*map(float, b)))
    g.add_segment(model.SegmentDef(na, nb, w, h, COPPER,
                                   None if width_dir is None else tuple(map(float, width_dir))))
    return model.pack(g, [], grid or model.Discretization.uniform(1, 1))


def filament(a, b, w, h, width_dir=None):
    return oracle.discretize(single(a, b, w, h, width_dir))[0]


# --------------------------------------------------------------------------
# SampledFilaments (neumann.rs:280-325)

def brute_mutual(a, b, panels, cross):
    """Sixfold composite Gauss-Legendre sum of the Neumann volume integral
    (tests/inductance.rs:92-118): ``panels`` panels of 8 nodes along each
    length, ``cross`` nodes across width and height."""
    x8, w8 = gl01(8)
    xc, wc = gl01(cross)

    def cloud(f):
        s = ((np.arange(panels)[:, None] + x8[None, :]) / panels - 0.5).ravel()
        ws = np.tile(w8, panels) * f["length"] / panels
        S, Y, Z = np.meshgrid(s, xc, xc, indexing="ij")
        W = ws[:, None, None] * wc[None, :, None] * wc[None, None, :]
        c = 0.5 * (f["start"] + f["end"])
        p = (c + S[..., None] * (f["length"] * f["dir"]) + (Y - 0.5)[..., None] * (f["width"] * f["wdir"])
             + (Z - 0.5)[..., None] * (f["height"] * f["hdir"]))
        return p.reshape(-1, 3), W.ravel()

    pa, wa = cloud(a)
    pb, wb = cloud(b)
    d = np.sqrt(((pa[:, None, :] - pb[None, :, :]) ** 2).sum(-1))
    return K * float(np.dot(a["dir"], b["dir"])) * float(wa @ (wb[None, :] / d).sum(1))


GENERAL_CASES = [
    # far: point quadrature
    (([0.5, 2.0, 1.0], [1.5, 2.8, 0.6], 0.08, 0.04, None), "point_quadrature"),
    (([3.0, 1.0, -2.0], [2.0, 2.0, -1.0], 0.2, 0.1, None), "point_quadrature"),
    # near, skew: sampled filaments
    (([0.2, 0.25, 0.1], [0.9, 0.5, 0.3], 0.06, 0.03, None), "sampled_filaments"),
    (([0.1, -0.2, 0.2], [0.8, 0.3, 0.25], 0.05, 0.05, None), "sampled_filaments"),
    # near, parallel, cross-section rotated by 30 degrees: sampled filaments
    (([0.1, 0.0, 0.3], [0.9, 0.0, 0.3], 0.08, 0.04, [0.0, 0.866, 0.5]), "sampled_filaments"),
]


@pytest.mark.parametrize("case", GENERAL_CASES, ids=lambda c: c[1])
def test_general_orientation_matches_a_fine_quadrature(case):
    """tests/inductance.rs:565-612"""
    (b0, b1, w, h, wd), method = case
    a = filament([0, 0, 0], [1, 0, 0], 0.10, 0.05)
    b = filament(b0, b1, w, h, wd)
    got, m, resolved = oracle.mutual(a, b)
    assert m == method and resolved
    want = brute_mutual(a, b, 6, 4)
    tol = 1e-7 if method == "point_quadrature" else 2e-6
    assert rel(got, want) < tol, (got, want)


def test_brute_force_reference_is_converged():
    """The sixfold reference itself is converged well below 2e-6 on the
    sampled cases: refining it changes it by far less."""
    a = filament([0, 0, 0], [1, 0, 0], 0.10, 0.05)
    for (b0, b1, w, h, wd), method in GENERAL_CASES[2:]:
        b = filament(b0, b1, w, h, wd)
        assert rel(brute_mutual(a, b, 6, 4), brute_mutual(a, b, 10, 6)) < 2e-7


def sampled_line_reference(a, b, na, nb):
    """The reference's finer sampling (tests/inductance.rs:663-690): the
    line-to-line formula over na^2 x nb^2 Gauss-sampled filaments."""
    ya, wa = gl01(na)
    yb, wb = gl01(nb)

    def lines(f, y, wy):
        out = []
        for yy, wyy in zip(y, wy):
            for zz, wzz in zip(y, wy):
                off = f["wdir"] * ((yy - 0.5) * f["width"]) + f["hdir"] * ((zz - 0.5) * f["height"])
                out.append((f["start"] + off, f["end"] + off, wyy * wzz))
        return out

    cos = float(np.dot(a["dir"], b["dir"]))
    total = 0.0
    for (a0, a1, wa_) in lines(a, ya, wa):
        for (b0, b1, wb_) in lines(b, yb, wb):
            v = oracle.lines_skew(a0, a1, b0, b1, 1e-6)
            assert v is not None
            total += wa_ * wb_ * v
    return K * cos * total


@pytest.mark.parametrize("case", [
    ("45 deg bend", [1.0, 0.0, 0.0], [1.7, 0.7, 0.0]),
    ("135 deg bend", [1.0, 0.0, 0.0], [0.3, 0.7, 0.0]),
    ("10 deg bend", [1.0, 0.0, 0.0], [1.98, 0.17, 0.0]),
    ("interpenetrating crossing", [0.2, -0.3, 0.0], [0.8, 0.3, 0.0]),
], ids=lambda c: c[0])
def test_touching_and_overlapping_skew_bars(case):
    """tests/inductance.rs:641-700: unresolved sampled filaments within 1e-3
    of a sampling twice as fine in every cross-section dimension."""
    _, b0, b1 = case
    a = filament([0, 0, 0], [1, 0, 0], 0.1, 0.05)
    b = filament(b0, b1, 0.1, 0.05)
    got, m, resolved = oracle.mutual(a, b)
    assert (m, resolved) == ("sampled_filaments", False)
    assert math.isfinite(got)
    want = sampled_line_reference(a, b, 16, 17)
    assert rel(got, want) < 1e-3, (got, want)


def test_sampled_filaments_parallel_limit():
    """tests/inductance.rs:706-759: close slender bars rotated out of
    alignment reach the bar closed form (Hoer-Love) within 1e-5, and tilts
    down to 2e-6 rad (the skew line formula) converge to it."""
    a = filament([0, 0, 0], [1, 0, 0], 0.01, 0.01, [0, 1, 0])
    cd = oracle.corner_differences
    v, est = oracle.bar_integral(cd(1, 1, 0), cd(0.01, 0.01, 0.1), cd(0.01, 0.01, 0))
    assert est < 1e-7
    exact = K * v / (0.01 ** 4)
    # the closed form against the sixfold reference: a separate check of the
    # value the limit is measured against
    b_aligned = filament([0, 0.1, 0], [1, 0.1, 0], 0.01, 0.01, [0, 1, 0])
    assert rel(brute_mutual(a, b_aligned, 16, 4), exact) < 1e-8
    rotated = filament([0, 0.1, 0], [1, 0.1, 0], 0.01, 0.01, [0, 1, 1])
    got, m, resolved = oracle.mutual(a, rotated)
    assert (m, resolved) == ("sampled_filaments", True)
    assert rel(got, exact) < 1e-5
    for angle in (1e-2, 1e-3, 1e-4, 1e-5, 2e-6):
        s, c = math.sin(angle), math.cos(angle)
        tilted = filament([0.5 - 0.5 * c, 0.1 - 0.5 * s, 0.0], [0.5 + 0.5 * c, 0.1 + 0.5 * s, 0.0],
                          0.01, 0.01, [-s, c, 0.0])
        got, m, _ = oracle.mutual(a, tilted)
        assert m == "sampled_filaments", angle
        assert rel(got, exact) < 20 * angle * angle + 1e-5, angle


# --------------------------------------------------------------------------
# Graded discretization (filament.rs:220-546, filament/grading.rs)

SKEW = dict(a=[0.3, -1.2, 2.0], b=[1.7, 0.4, -0.5], w=0.6, h=0.25, width_dir=[0.2, 1.0, 0.1])
TOL = 1e-12


def skew_graded(nw, nh, rw, rh=None):
    rh = rw if rh is None else rh
    grid = model.Discretization.per_segment_graded([model.AxisGrading(nw, nh, rw, rh)])
    return oracle.discretize(single(SKEW["a"], SKEW["b"], SKEW["w"], SKEW["h"],
                                    SKEW["width_dir"], grid))


def skew_uniform(nw, nh):
    return oracle.discretize(single(SKEW["a"], SKEW["b"], SKEW["w"], SKEW["h"], SKEW["width_dir"],
                                    model.Discretization.uniform(nw, nh)))


def centres(f):
    return 0.5 * (f["start"] + f["end"])


def geometric_extents(total, count, ratio):
    """The graded partition from its definition: weights ratio^min(i, n-1-i),
    normalised by their closed-form sum."""
    k = np.minimum(np.arange(count), count - 1 - np.arange(count))
    weights = ratio ** k.astype(float)
    half = count // 2
    if ratio == 1.0:
        s = float(count)
    else:  # two geometric series (plus the middle term of an odd count)
        s = 2.0 * (ratio ** half - 1.0) / (ratio - 1.0) + (ratio ** half if count % 2 else 0.0)
    return total * weights / s


def test_graded_surface_extent_doc_values():
    """filament.rs:243-258 and :1044-1052."""
    assert model.graded_surface_extent(8.0, 4, 3.0) == 1.0
    assert abs(model.graded_surface_extent(8.0, 6, 3.0) - 8.0 / 26.0) < 1e-15
    assert abs(model.graded_surface_extent(1e-3, 16, 1.0) - 62.5e-6) < 1e-18
    g = model.graded_surface_extent(1e-3, 6, 4.0)
    assert abs(g - 1e-3 / (1 + 4 + 16 + 16 + 4 + 1)) < 1e-18 and g < 0.4 * 62.5e-6


def test_graded_grid_matches_hand_computed_fixture():
    """filament.rs:1020-1038: weights 1, 3, 3, 1 over a 0.8-wide segment."""
    f = oracle.discretize(single([0, 0, 0], [1, 0, 0], 0.8, 0.2, None,
                                 model.Discretization.graded(4, 1, 3.0)))
    want = [(-0.35, 0.1), (-0.15, 0.3), (0.15, 0.3), (0.35, 0.1)]
    assert len(f) == 4
    for fi, (y, w) in zip(f, want):
        assert np.allclose(fi["start"], [0.0, y, 0.0], atol=TOL, rtol=0)
        assert np.allclose(fi["end"], [1.0, y, 0.0], atol=TOL, rtol=0)
        assert abs(fi["width"] - w) < TOL and abs(fi["height"] - 0.2) < TOL


@pytest.mark.parametrize("nw,nh,ratio", [(4, 4, 10.0), (3, 5, 2.0), (1, 8, 1.5), (7, 1, 3.0)])
def test_graded_filaments_tile_the_cross_section(nw, nh, ratio):
    """filament.rs:935-953"""
    f = skew_graded(nw, nh, ratio)
    assert len(f) == nw * nh
    area = SKEW["w"] * SKEW["h"]
    assert abs(np.sum(f["width"] * f["height"]) - area) < TOL * area
    for i in range(len(f)):
        assert abs(f["width"][i] - f["width"][i % nw]) < TOL
        assert abs(f["height"][i] - f["height"][(i // nw) * nw]) < TOL


def test_graded_extents_follow_the_geometric_progression():
    """filament.rs:957-981, against the closed-form partition."""
    nw, nh, ratio = 6, 5, 2.5
    f = skew_graded(nw, nh, ratio)
    widths, heights = f["width"][:nw], f["height"][::nw]
    for ext, total in ((widths, SKEW["w"]), (heights, SKEW["h"])):
        n = len(ext)
        assert np.allclose(ext, geometric_extents(total, n, ratio), rtol=1e-14, atol=0)
        surface = model.graded_surface_extent(total, n, ratio)
        assert abs(ext[0] - surface) < TOL * surface and abs(ext[-1] - surface) < TOL * surface
        for i in range(n // 2):
            assert abs(ext[i] - ext[n - 1 - i]) < TOL * ext[i]
            if i + 1 < n // 2:
                assert abs(ext[i + 1] - ratio * ext[i]) < TOL * ratio * ext[i]


@pytest.mark.parametrize("nw,nh,ratio", [(4, 4, 10.0), (5, 3, 2.0), (2, 6, 1.25)])
def test_graded_bundle_is_symmetric_and_fills_the_surfaces(nw, nh, ratio):
    """filament.rs:985-1015"""
    f = oracle.discretize(single([0, 0, 0], [2, 0, 0], 0.6, 0.4, None,
                                 model.Discretization.graded(nw, nh, ratio)))
    c = centres(f)
    assert np.linalg.norm(c.mean(0) - [1.0, 0.0, 0.0]) < TOL
    assert abs(np.min(c[:, 1] - 0.5 * f["width"]) + 0.3) < TOL
    assert abs(np.max(c[:, 1] + 0.5 * f["width"]) - 0.3) < TOL
    assert abs(np.min(c[:, 2] - 0.5 * f["height"]) + 0.2) < TOL
    assert abs(np.max(c[:, 2] + 0.5 * f["height"]) - 0.2) < TOL


def rotation(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    x, y, z = axis
    c, s = math.cos(angle), math.sin(angle)
    C = 1 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def test_grading_commutes_with_rigid_motion():
    """filament.rs:1042-1068"""
    R = rotation([0.4, -1.0, 0.7], 1.234)
    shift = np.array([-2.0, 0.5, 3.0])
    a, b, wd = (np.array(SKEW[k], float) for k in ("a", "b", "width_dir"))
    grid = model.Discretization.graded(4, 3, 2.0)
    before = oracle.discretize(single(a, b, SKEW["w"], SKEW["h"], wd, grid))
    after = oracle.discretize(single(R @ a + shift, R @ b + shift, SKEW["w"], SKEW["h"], R @ wd, grid))
    assert len(after) == 12
    for f, g in zip(before, after):
        assert np.linalg.norm(R @ f["start"] + shift - g["start"]) < TOL
        assert np.linalg.norm(R @ f["end"] + shift - g["end"]) < TOL
        assert abs(f["width"] - g["width"]) < TOL and abs(f["height"] - g["height"]) < TOL


def test_unit_ratio_is_bit_identical_to_the_uniform_grid():
    """filament.rs:872-884"""
    for nw, nh in [(1, 1), (1, 4), (3, 1), (3, 2), (5, 7)]:
        assert skew_graded(nw, nh, 1.0).tobytes() == skew_uniform(nw, nh).tobytes()


@pytest.mark.parametrize("nw,nh,rw,rh", [(4, 3, 3.0, 1.0), (3, 5, 1.0, 2.0), (5, 4, 2.0, 1.5)])
def test_per_axis_ratios_grade_each_axis_independently(nw, nh, rw, rh):
    """filament.rs:886-933"""
    mixed = skew_graded(nw, nh, rw, rh)
    by_w = skew_graded(nw, nh, rw)
    by_h = skew_graded(nw, nh, rh)
    widths = geometric_extents(SKEW["w"], nw, rw)
    heights = geometric_extents(SKEW["h"], nh, rh)
    mid = 0.5 * (np.array(SKEW["a"]) + np.array(SKEW["b"]))
    for idx in range(nw * nh):
        i, j = idx % nw, idx // nw
        f = mixed[idx]
        assert abs(f["width"] - widths[i]) < TOL and abs(f["height"] - heights[j]) < TOL
        off = centres(mixed)[idx] - mid
        assert abs(off @ f["wdir"] - (centres(by_w)[idx] - mid) @ f["wdir"]) < TOL
        assert abs(off @ f["hdir"] - (centres(by_h)[idx] - mid) @ f["hdir"]) < TOL
    area = SKEW["w"] * SKEW["h"]
    assert abs(np.sum(mixed["width"] * mixed["height"]) - area) < TOL * area
    assert skew_graded(3, 4, 2.5, 2.5).tobytes() == skew_graded(3, 4, 2.5).tobytes()


@pytest.mark.parametrize("ratio", [0.5, 0.0, -2.0, math.inf, math.nan])
def test_invalid_grading_is_rejected(ratio):
    """filament.rs:1070-1093: the host rejects it when resolving, and the
    oracle and the product reject a packed grid that carries it anyway
    (the oracle as DiscretizeError::InvalidRatio, sub-code 12; the product
    as the SolveError::InvalidGrading that assembly reports, solve.rs:477-489)."""
    with pytest.raises(InvalidGrading):
        model.Discretization.graded(2, 2, ratio).resolve(
            model.Geometry.from_parts([model.Node(0, 0, 0), model.Node(1, 0, 0)],
                                      [model.SegmentDef(0, 1, 0.1, 0.1, COPPER)]))
    packed = single(SKEW["a"], SKEW["b"], SKEW["w"], SKEW["h"], SKEW["width_dir"],
                    model.Discretization.graded(2, 2, 2.0))
    packed.grids["width_ratio"] = ratio
    with pytest.raises(oracle.OracleError) as e:
        oracle.discretize(packed)
    assert e.value.detail.code == 12
    with pytest.raises(InvalidGrading):
        discretize_packed(packed)


def test_zero_subdivision_is_rejected():
    for nw, nh in [(0, 1), (1, 0), (0, 0)]:
        packed = single(SKEW["a"], SKEW["b"], SKEW["w"], SKEW["h"], SKEW["width_dir"],
                        model.Discretization.graded(2, 2, 2.0))
        packed.grids["nw"] = nw
        packed.grids["nh"] = nh
        with pytest.raises(oracle.OracleError) as e:
            oracle.discretize(packed)
        assert e.value.detail.code == 10


# --------------------------------------------------------------------------
# SkinDepthGrading::count_for (solve.rs:388-402)

def brute_count(extent, sigma, frequency, target_depths, ratio, cap):
    delta = 1.0 / math.sqrt(math.pi * frequency * 4e-7 * math.pi * sigma) if frequency > 0 else math.inf
    target = target_depths * delta
    if math.isinf(target) or extent <= target:
        return 1
    for n in range(2, cap + 1):
        if geometric_extents(extent, n, ratio)[0] <= target * (1 + 1e-15):
            return n
    return cap


def test_skin_depth_formula():
    """solve.rs:560-562, tests/currents.rs:92: delta = 1/sqrt(pi f mu0 sigma)."""
    f = 1e9
    want = math.sqrt(2.0 / (2 * math.pi * f * 4e-7 * math.pi * COPPER))
    assert abs(model.skin_depth(f, COPPER) - want) < 1e-12 * want
    assert model.skin_depth(0.0, COPPER) == math.inf


def test_count_for_matches_brute_force():
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(3000):
        extent = 10 ** rng.uniform(-6, -2.5)
        freq = 10 ** rng.uniform(3, 11)
        target = rng.choice([0.25, 0.5, 1.0, 2.0])
        ratio = rng.choice([1.0, 1.25, 2.0, 3.0, 4.0])
        cap = int(rng.choice([4, 16, 32]))
        g = model.SkinDepthGrading(freq, float(target), float(ratio), cap)
        got = g.count_for(extent, COPPER)
        want = brute_count(extent, COPPER, freq, target, ratio, cap)
        if got != want:
            # only a tie at the rounding of the surface extent may differ
            surf = geometric_extents(extent, got, ratio)[0]
            assert abs(surf / (target * model.skin_depth(freq, COPPER)) - 1) < 1e-12
        checked += got > 1
    assert checked > 1000


def test_count_for_reference_cases():
    """tests/skin_effect_validation.rs:391-413 (DC is one filament; counts do
    not fall with frequency and never exceed the uniform grid's) and
    tests/graded_grid_validity.rs:88-98 (an underflowed skin depth saturates
    the cap instead of falling back to one filament)."""
    width, thickness = 2e-3, 35e-6
    dc = model.SkinDepthGrading(0.0, 0.5, 2.0)
    assert dc.count_for(width, COPPER) == 1 and dc.count_for(thickness, COPPER) == 1
    previous = 0
    for per_delta in (0.5, 2.0, 10.0, 100.0):
        delta = thickness / per_delta
        f = 1.0 / (math.pi * 4e-7 * math.pi * COPPER * delta * delta)
        g = model.SkinDepthGrading(f, 0.5, 2.0)
        count = g.count_for(width, COPPER) * g.count_for(thickness, COPPER)
        uniform = math.ceil(width / (0.5 * delta)) * math.ceil(thickness / (0.5 * delta))
        assert previous <= count <= uniform
        previous = count
    g = model.SkinDepthGrading(1e308, 0.5, 4.0)
    assert g.count_for(1e-3, COPPER) == 32


# --------------------------------------------------------------------------
# The product's discretization (fh_geometry.cpp) against the oracle's, bit
# for bit, on the graded grids the batch and adjoint workloads run on.  Host
# code: it runs in the CPU suite, and once more on the GPU box.

def graded_problems():
    out = {
        "graded_3x3_r2": fixtures.graded_spiral(3, 100e-6, 50e-6, 35e-6, 3, 3, 2, 2.0),
        "per_axis": fixtures.graded_busbar_pair(200e-6, 100e-6, 35e-6, 5, 3, 4, ratio=3.0),
        "adjoint_spiral": fixtures.AdjointSpiral().problem(),
        "skin_depth_meander": (lambda g, p: (g, [p], model.Discretization.skin_depth(1e9, 0.5, 2.0)))(
            *fixtures.meander(6)),
    }
    for k in (0, 1, 2, 3, 20, 47):
        out[f"batch{k}"] = fixtures.batch_problems()[k]
    return out


@pytest.mark.parametrize("suite", ["host", pytest.param("device_box", marks=pytest.mark.gpu)])
@pytest.mark.parametrize("name", list(graded_problems().keys()))
def test_product_discretization_is_bit_identical_on_graded_grids(name, suite):
    g, ports, d = graded_problems()[name]
    if name == "per_axis":
        n = g.segment_count()
        d = model.Discretization.per_segment_graded([model.AxisGrading(5, 3, 3.0, 1.5)] * n)
    packed = model.pack(g, ports, d)
    assert packed.grids["graded"].all() and (packed.grids["width_ratio"] > 1).all()
    got = discretize_packed(packed)
    want = oracle.discretize(packed)
    assert len(got) == len(want) == packed.filament_count
    assert got.tobytes() == want.tobytes()
