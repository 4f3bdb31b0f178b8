"""Initial suspensions (mds_init_radii / mds_init_positions /
mds_init_orientations): host code, no GPU. The cell-list placement is pinned
against a brute-force O(N^2) replay of the same random stream
(src/simulation.cpp:238-265 rule), and the samplers against the
distributions they must draw from."""
import numpy as np
import pytest

import md_sim_wolfschwenger_amd as mds

M64 = (1 << 64) - 1


class Xoshiro:
    """xoshiro256** seeded by splitmix64, restated from the published
    algorithm (Blackman and Vigna, ACM TOMS 47, 2021)."""

    def __init__(self, seed):
        z = seed & M64
        self.s = []
        for _ in range(4):
            z = (z + 0x9E3779B97F4A7C15) & M64
            x = z
            x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
            x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
            self.s.append(x ^ (x >> 31))

    @staticmethod
    def rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & M64

    def next(self):
        s = self.s
        r = (self.rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self.rotl(s[3], 45)
        return r

    def balanced(self):
        return 2.0 * ((self.next() >> 11) * 2.0 ** -53) - 1.0


def cfg_of(text=""):
    return mds.config_from_string("[initial_configuration]\nINIT_CONFIG_TRANS = 1\nINIT_CONFIG_ROT = 2\n" + text)


@pytest.mark.parametrize("periodic", [0, 1])
def test_placement_equals_bruteforce_replay(periodic):
    cfg = cfg_of()
    rng = np.random.default_rng(4)
    n = 150
    rh = 1.0 + 0.5 * rng.random(n)
    L = (np.sum(4 / 3 * np.pi * rh ** 3) / 0.25) ** (1 / 3)  # 25 % volume fraction
    seed = 987654321
    pos = mds.init_positions(cfg, rh, L, seed, periodic=bool(periodic))
    g = Xoshiro(seed)
    ref = np.zeros((n, 3))
    for i in range(n):
        ext = L / 2 if periodic else L / 2 - rh[i]
        while True:
            p = np.array([g.balanced() * ext for _ in range(3)])
            d = ref[:i] - p
            if periodic:
                d -= L * np.round(d / L)
            if i == 0 or np.all(np.einsum("ij,ij->i", d, d) >= (rh[i] + rh[:i]) ** 2):
                break
        ref[i] = p
    assert np.array_equal(pos, ref)


def test_placement_bounds_and_no_overlap_large():
    cfg = cfg_of()
    n = 20000
    r = np.full(n, 1e-8)
    rh = r + 5e-9
    L = (np.sum(4 / 3 * np.pi * rh ** 3) / 0.10) ** (1 / 3)
    pos = mds.init_positions(cfg, rh, L, 5, periodic=False)
    assert np.all(np.abs(pos) <= (L / 2 - rh)[:, None])
    from scipy.spatial import cKDTree
    tr = cKDTree(pos)
    pairs = np.array(sorted(tr.query_pairs(2 * rh.max() * (1 - 1e-12))))
    assert len(pairs) == 0 or np.all(
        np.linalg.norm(pos[pairs[:, 0]] - pos[pairs[:, 1]], axis=1) >= rh[pairs[:, 0]] + rh[pairs[:, 1]])


def test_grid_positions():
    cfg = mds.config_from_string("[initial_configuration]\nINIT_CONFIG_TRANS = 2\n[particle_properties]\nnpDim = 3\n")
    pos = mds.init_positions(cfg, np.full(27, 1.0), 9.0, 0)
    ip = 0
    for ix in range(3):
        for iy in range(3):
            for iz in range(3):
                assert np.allclose(pos[ip], [(ix + 0.5) * 3 - 4.5, (iy + 0.5) * 3 - 4.5, (iz + 0.5) * 3 - 4.5])
                ip += 1


def test_radii_distributions():
    n = 200000
    c = cfg_of("[assemble_properties]\nrMagMean = 1e-8\n")
    assert np.all(mds.init_radii(c, 10, 1) == 1e-8)
    c = cfg_of("[size_distribution]\nSIZE_DIST = 1\n[assemble_properties]\nrMagMean = 1e-8\n"
               "[mi355x]\nrMagMedian = 1.2e-8\nsigmaLog = 0.25\n")
    lr = np.log(mds.init_radii(c, n, 2))
    assert abs(np.median(lr) - np.log(1.2e-8)) < 4 * 0.25 * 1.25 / np.sqrt(n)
    assert abs(lr.std() - 0.25) < 0.003
    # SIZE_DIST_LOG by mean and standard deviation (sigma = derived from the ini)
    c = cfg_of("[size_distribution]\nSIZE_DIST = 1\n[assemble_properties]\nrMagMean = 1e-8\n")
    if c.sigma > 0:
        r = mds.init_radii(c, n, 3)
        assert abs(r.mean() - 1e-8) < 5 * c.sigma / np.sqrt(n)
        assert abs(r.std() - c.sigma) < 0.02 * c.sigma
    assert not np.array_equal(mds.init_radii(c, 100, 1) if c.sigma > 0 else np.zeros(1),
                              mds.init_radii(c, 100, 2) if c.sigma > 0 else np.ones(1))


@pytest.mark.parametrize("rot", [1, 2, 3])
def test_orientations(rot):
    c = mds.config_from_string(f"[initial_configuration]\nINIT_CONFIG_ROT = {rot}\n"
                               "[physical_constants]\nmy_pi = 3.14159265359\n")
    n = 100000
    mm, ea = mds.init_orientations(c, n, 9)
    assert np.allclose(np.linalg.norm(ea, axis=1), 1.0)
    if rot == 3:
        assert np.all(ea == [0, 0, 1]) and np.all(mm == ea)
        return
    # uniform on the sphere: mean 0, <e_d^2> = 1/3
    assert np.all(np.abs(ea.mean(axis=0)) < 5 / np.sqrt(3 * n))
    assert np.all(np.abs((ea ** 2).mean(axis=0) - 1 / 3) < 0.01)
    if rot == 1:
        assert np.all(mm == [0, 0, 1])
    else:
        assert np.array_equal(mm, ea)
