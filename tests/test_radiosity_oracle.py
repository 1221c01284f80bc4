"""CPU pins of the radiative-heat restatement (RSolverRadiativeHeat with a loaded view-factor
matrix: patch means, the dense net-radiation system, the area split of the patch heat) against
textbook closed forms of gray-body exchange:

* a gray patch facing black surroundings only (F = 0): q = E sigma (Tamb^4 - T^4);
* two infinite parallel plates (F12 = F21 = 1): q1 = -sigma (T1^4 - T2^4) / (1/E1 + 1/E2 - 1),
  with each plate made of elements of different emissivity and temperature (patch means);
* a closed enclosure with reciprocity (A_i F_ij = A_j F_ji, rows summing to 1): the net heats
  balance, sum_i A_i q_i = 0, and a uniform temperature gives q = 0;
* a patch of zero emissivity contributes a zero column;
* the element heats of a patch sum to the patch heat in proportion to the element areas.
"""
import numpy as np

SIGMA = 5.670367e-8


def one_elem_patches(n):
    return np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32)


def test_gray_patch_to_black_surroundings(orc):
    pp, pe = one_elem_patches(1)
    for E, T, Ta in ((0.7, 600.0, 300.0), (0.2, 300.0, 900.0)):
        A, b = orc.radiosity_assemble(np.zeros((1, 1)), pp, pe, [E], [T], [Ta])
        q = np.linalg.solve(A, b)
        np.testing.assert_allclose(q, [E * SIGMA * (Ta ** 4 - T ** 4)], rtol=1e-13)


def test_parallel_plates_with_patch_means(orc):
    # plate 1 = elements 0, 1; plate 2 = elements 2, 3, 4
    pp = np.array([0, 2, 5], np.uint32)
    pe = np.array([1, 0, 4, 2, 3], np.uint32)
    em = np.array([0.4, 0.6, 0.9, 0.75, 0.6])
    te = np.array([1000.0, 1200.0, 300.0, 500.0, 400.0])
    am = np.full(5, 250.0)
    F = np.array([[0.0, 1.0], [1.0, 0.0]])
    A, b = orc.radiosity_assemble(F, pp, pe, em, te, am)
    q = np.linalg.solve(A, b)
    E1, E2, T1, T2 = 0.5, 0.75, 1100.0, 400.0
    q1 = -SIGMA * (T1 ** 4 - T2 ** 4) / (1 / E1 + 1 / E2 - 1)
    np.testing.assert_allclose(q, [q1, -q1], rtol=1e-12)


def enclosure(rng, n):
    G = rng.uniform(0.0, 1.0, (n, n))
    G = G + G.T
    np.fill_diagonal(G, 0.0)
    area = G.sum(1)
    return G / area[:, None], area


def test_enclosure_energy_balance(orc):
    rng = np.random.default_rng(2)
    n = 40
    F, area = enclosure(rng, n)
    pp, pe = one_elem_patches(n)
    E = rng.uniform(0.1, 0.95, n)
    T = rng.uniform(300, 1500, n)
    A, b = orc.radiosity_assemble(F, pp, pe, E, T, rng.uniform(200, 400, n))
    q = np.linalg.solve(A, b)
    assert abs(area @ q) <= 1e-10 * (area @ np.abs(q))
    assert np.abs(q).max() > 1.0
    A, b = orc.radiosity_assemble(F, pp, pe, E, np.full(n, 777.0), np.full(n, 300.0))
    assert np.abs(np.linalg.solve(A, b)).max() <= 1e-9 * SIGMA * 777.0 ** 4


def test_zero_emissivity_column(orc):
    rng = np.random.default_rng(3)
    n = 6
    F, _ = enclosure(rng, n)
    pp, pe = one_elem_patches(n)
    E = rng.uniform(0.2, 0.9, n)
    E[2] = 0.0
    A, _ = orc.radiosity_assemble(F, pp, pe, E, rng.uniform(300, 900, n), np.full(n, 300.0))
    assert np.all(A[:, 2] == 0.0)
    assert np.all(A[:, [0, 1, 3, 4, 5]] != 0.0)


def test_patch_heat_split_by_area(orc):
    pp = np.array([0, 3, 3, 5], np.uint32)          # patch 1 has no elements
    pe = np.array([4, 0, 2, 5, 1], np.uint32)
    area = np.array([1.0, 2.5, 0.5, 9.0, 3.0, 1.5])  # element 3 is in no patch
    heat = np.array([12.0, -7.0, 5.0])
    out = orc.radiosity_process(pp, pe, area, heat)
    np.testing.assert_allclose(out, [12.0 * 1.0 / 4.5, 5.0 * 2.5 / 4.0, 12.0 * 0.5 / 4.5, 0.0, 12.0 * 3.0 / 4.5,
                                     5.0 * 1.5 / 4.0], rtol=1e-15)
