"""CPU checks that pin the oracle's direct-light estimator
(orc_gbuffer_cast_direct): the environment sampling tables against
content-independent properties, the BSDF's reciprocity and the diffuse
lobe's normalisation, and the estimator's mean over many samples against an
independent quadrature of its expectation.

The estimator restates the closest-hit program's direct term in wavelet
mode (Renderer/device_functions/devicePrograms.cu:165-280) with the
cookTorrance and roughPlastic BSDFs
(Renderer/device_functions/cookTorrance.cu:26-195, roughPlastic.cu:25-174)
and the environment tables of Renderer/EnvironmentMap.cpp:128-190 and
Renderer/table_dist.cpp:4-130.  Its pdf is per unit face (u, v) area
(cubeMap.cu:9-19), not per solid angle, so its expectation is the
quadrature below, not the rendering-equation integral."""
import numpy as np
import pytest

from oracle import binding as orc

LUM = np.array([0.212671, 0.715160, 0.072169])


def _lum(x):
    return np.asarray(x, np.float64) @ LUM


def test_tables_sum_to_one_and_face_pmf_follows_luminance():
    rng = np.random.default_rng(0)
    faces = rng.lognormal(0, 1, (6, 8, 8, 3)).astype(np.float32)
    faces[3] *= 5
    t = orc.env_tables(faces)
    lum = np.array([_lum(faces[f].reshape(-1, 3)).sum() for f in range(6)])
    assert np.allclose(t["face_pmf"], lum / lum.sum(), rtol=1e-6)
    assert np.allclose(np.cumsum(np.r_[0, t["face_pmf"]])[:6], t["face_cdf"][:6], atol=1e-15)
    assert np.allclose(t["pdf_rows"].sum(axis=2), 1.0, rtol=1e-12)
    assert np.allclose(t["pdf_marginals"].sum(axis=1), 1.0, rtol=1e-12)
    assert np.all(np.diff(t["cdf_rows"], axis=2) >= 0)
    assert np.all(t["cdf_rows"][:, :, -1] == 1.0) and np.all(t["cdf_marginals"][:, -1] == 1.0)


def test_table_cell_is_the_texel_it_weights():
    # Doubling texel (r, c) of face f changes only row r of that face, and
    # within it raises column c's share from w / S to 2 w / (S + w).
    n = 8
    base = np.ones((6, n, n, 3), np.float32)
    t0 = orc.env_tables(base)
    for f, r, c in ((0, 1, 6), (4, 7, 0), (2, 3, 3)):
        x = base.copy()
        x[f, r, c] *= 2
        t1 = orc.env_tables(x)
        d = np.abs(t1["pdf_rows"] - t0["pdf_rows"]) > 1e-15
        assert d[f, r].any() and d.sum() == d[f, r].sum()
        p = t0["pdf_rows"][f, r, c]
        assert t1["pdf_rows"][f, r, c] == pytest.approx(2 * p / (1 + p), rel=1e-6)
    # The weights do not depend on the face.
    assert np.allclose(t0["pdf_rows"], t0["pdf_rows"][0][None], rtol=1e-12)


MAT = [0.7, 0.5, 0.3, 0.0, 0.0, 0.0, 0.4, 2]
UP = np.array([0.0, 1.0, 0.0])


def _face_dir(f, u, v):
    uc, vc = 2 * u - 1, 2 * v - 1
    one = np.ones_like(uc)
    d = {0: (one, vc, uc), 1: (-one, vc, -uc), 2: (uc, one, vc), 3: (uc, -one, -vc),
         4: (uc, vc, -one), 5: (-uc, vc, one)}[f]
    d = np.stack(d, -1)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def _bilinear(face, u, v):
    n = face.shape[0]
    x, y = u * n - 0.5, v * n - 0.5
    x0, y0 = np.floor(x).astype(int), np.floor(y).astype(int)
    wx, wy = x - x0, y - y0
    x1, y1 = (x0 + 1) % n, (y0 + 1) % n
    x0, y0 = x0 % n, y0 % n
    top = face[y0, x0] * (1 - wx)[..., None] + face[y0, x1] * wx[..., None]
    bot = face[y1, x0] * (1 - wx)[..., None] + face[y1, x1] * wx[..., None]
    return top * (1 - wy)[..., None] + bot * wy[..., None]


def _grid(n, s):
    m = n * s
    g = (np.arange(m) + 0.5) / m
    U, V = np.meshgrid(g, g)
    J = 4.0 / (1 + (2 * U - 1) ** 2 + (2 * V - 1) ** 2) ** 1.5  # d(omega) / d(u, v)
    return U.ravel(), V.ravel(), J.ravel(), 1.0 / (m * m)


@pytest.mark.parametrize("kind", [0, 2, 4])  # roughPlastic, cookTorrance, simple_ggx
def test_bsdf_reciprocity_and_diffuse_pdf_normalisation(kind):
    rng = np.random.default_rng(3)
    mat = [0.6, 0.4, 0.2, 0.3, 0.3, 0.3, 0.35, kind]
    for _ in range(20):
        a, b = rng.normal(size=3), rng.normal(size=3)
        a[1], b[1] = abs(a[1]) + 0.05, abs(b[1]) + 0.05
        a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
        fab = orc.ct_eval(mat, UP, UP, a, b) / a[1]   # f(a <- b)
        fba = orc.ct_eval(mat, UP, UP, b, a) / b[1]
        assert np.allclose(fab, fba, rtol=1e-9)
    if kind == 4:
        return  # simple_ggx always keeps a specular lobe (probability >= 0.05)
    # With ks = 0 the sampler is the cosine lobe; its pdf integrates to 1.
    U, V, J, du = _grid(8, 6)
    v = np.array([0.3, 0.8, 0.2])
    v /= np.linalg.norm(v)
    tot = 0.0
    diffuse = MAT[:7] + [kind]
    for f in range(6):
        for d, j in zip(_face_dir(f, U, V), J):
            tot += orc.ct_pdf(diffuse, UP, UP, d, v) * j * du
    assert tot == pytest.approx(1.0, abs=2e-3)


@pytest.mark.parametrize("kind", [0, 2])  # roughPlastic, cookTorrance
def test_estimator_mean_matches_quadrature_of_its_expectation(kind):
    """Diffuse roughPlastic / cookTorrance (ks = 0: cosine sampling, pdf
    exact) on an open ground plane (nothing above it to shadow), smooth
    environment.
    E = integral over faces of pe^2 g / max(pe^2 + pb^2, 1e-8) d(u, v)
      + integral of pb^2 g / max(pe^2 + pb^2, 1e-8) d(omega), pe > 1e-4,
    with g = BSDF x cosine x radiance; the mean of 8 x 2000 replayed
    samples must lie within 5 standard errors (+ 1e-3) of it."""
    n = 8
    mat = MAT[:7] + [kind]
    rng = np.random.default_rng(11)
    light = rng.lognormal(0, 0.5, (6, n, n, 3)).astype(np.float32)
    ground = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    ix = np.array([[0, 2, 1], [0, 3, 2]], np.int32)  # CCW from +y: normal +y
    cam = [0.0, 2.0, 3.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0]
    means = []
    for seed in range(8):
        g = orc.gbuffer_cast(ground, ix, [mat], cam, 2, 2, bounds=[-60, -1, -60, 60, 1, 60],
                             light_faces=light, samples=2000, seed=seed)
        assert g["env_flag"][3] == 1
        means.append(g["direct"][3, :3])
    means = np.array(means)
    est, sem = means.mean(0), means.std(0, ddof=1) / np.sqrt(len(means))
    t = orc.env_tables(light)
    view = -np.array(cam[3:6]) + np.array(cam[0:3])
    view /= np.linalg.norm(view)
    U, V, J, du = _grid(n, 8)
    e = np.zeros(3)
    for f in range(6):
        dirs = _face_dir(f, U, V)
        rad = _bilinear(light[f].astype(np.float64), U, V)
        for d, j, le in zip(dirs, J, rad):
            pe = orc.env_pdf(t, n, d)
            if pe <= 1e-4:
                continue
            gv = orc.ct_eval(mat, UP, UP, d, view) * le
            if not gv.any():
                continue
            pb = orc.ct_pdf(mat, UP, UP, d, view)
            den = max(pe * pe + pb * pb, 1e-8)
            e += (pe * pe * gv / den + pb * pb * gv / den * j) * du
    assert np.all(np.abs(est - e) <= 5 * sem + 1e-3 * e), (est, e, sem)
