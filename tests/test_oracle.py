"""Pins of the CPU oracle (CPU only).

The reference ships no golden vectors for this path (its only known answer,
the H2O integration test, needs network data), so the oracle is pinned by
mathematics that a plausible transcription mistake breaks:

* tuple enumeration and closed-form decode: brute force, exhaustive;
* W: the simultaneous-permutation symmetry W^{abc}_{ijk} = W^{bac}_{jik} =
  W^{acb}_{ikj}, which holds for Tabij = Tbaji and fails if any of the twelve
  terms uses a wrong operand index or sign;
* E(T): the restricted sum over a<=b<=c and i>=j>=k with the distinct/same
  formulas equals one sixth of the unrestricted closed-shell sum
  sum_{abc,ijk} 2 W (4Z_ijk + Z_jki + Z_kij - 2Z_ikj - 2Z_jik - 2Z_kji) / D
  (W, Z built here with numpy from the tensor definitions), for inputs with
  Tabij = Tbaji and Vabij = Vbaji;
* planted mistakes in the oracle (one per term family, the weights, the
  formula selector, the singles indices) each break that identity;
* invariants: eps shift, V = 0, scaling E(lT, mV) = l^2 m^2 E at Tai = 0,
  ijkabc sign, sub-range additivity, (cT) with J = V.
"""
import numpy as np
import pytest


def _full_space_energy(no, nv, t):
    """-(1/6) sum over all (a,b,c) and (i,j,k) of the closed-shell kernel."""
    T = t["Tabij"].reshape((nv, nv, no, no), order="F")
    Vab = t["Vabij"].reshape((nv, nv, no, no), order="F")
    Vh = t["Vijka"].reshape((no, no, no, nv), order="F")
    Vp = t["Vabci"].reshape((nv, nv, nv, no), order="F")
    tai = t["Tai"].reshape((nv, no), order="F")
    ei, ea = t["ei"], t["ea"]
    es = np.einsum
    W = (es("aEij,bcEk->abcijk", T, Vp) + es("aEik,cbEj->abcijk", T, Vp)
         + es("cEki,abEj->abcijk", T, Vp) + es("cEkj,baEi->abcijk", T, Vp)
         + es("bEji,acEk->abcijk", T, Vp) + es("bEjk,caEi->abcijk", T, Vp)
         - es("abLj,ikLc->abcijk", T, Vh) - es("abiL,jkLc->abcijk", T, Vh)
         - es("acLk,ijLb->abcijk", T, Vh) - es("aciL,kjLb->abcijk", T, Vh)
         - es("bcLk,jiLa->abcijk", T, Vh) - es("bcjL,kiLa->abcijk", T, Vh))
    S = (es("ai,bcjk->abcijk", tai, Vab) + es("bj,acik->abcijk", tai, Vab)
         + es("ck,abij->abcijk", tai, Vab))
    Z = W + S
    tr = lambda X, p: X.transpose(0, 1, 2, *[3 + q for q in p])
    Y = (4 * Z + tr(Z, (1, 2, 0)) + tr(Z, (2, 0, 1)) - 2 * tr(Z, (0, 2, 1))
         - 2 * tr(Z, (1, 0, 2)) - 2 * tr(Z, (2, 1, 0)))
    D = (ea[:, None, None, None, None, None] + ea[None, :, None, None, None, None]
         + ea[None, None, :, None, None, None] - ei[None, None, None, :, None, None]
         - ei[None, None, None, None, :, None] - ei[None, None, None, None, None, :])
    return -np.sum(2 * W * Y / D) / 6.0, W


SIZES = [(1, 2), (2, 3), (3, 5), (4, 7), (5, 6), (3, 17)]


@pytest.mark.parametrize("nv", list(range(1, 41)) + [57, 64])
def test_decode_matches_enumeration(oracle, nv):
    allt = oracle.all_tuples(nv)
    assert len(allt) == oracle.n_tuples(nv) == nv * (nv + 1) * (nv + 2) // 6 - nv
    brute = [(a, b, c) for a in range(nv) for b in range(a, nv) for c in range(b, nv)
             if not (a == b == c)]
    assert [tuple(map(int, x)) for x in allt] == brute
    step = 1 if nv <= 40 else 7
    for it in range(0, len(brute), step):
        assert oracle.nth_tuple(it, nv) == brute[it]


@pytest.mark.parametrize("no,nv", SIZES)
def test_full_space_identity(oracle, no, nv):
    t = oracle.synthetic(no, nv, seed=3)
    e, ect = oracle.run(no, nv, t)
    e_full, _ = _full_space_energy(no, nv, t)
    assert ect == e  # no J tensors: (cT) accumulates the same energy
    # No = 1 has E = 0 in exact arithmetic: allow a roundoff floor
    assert abs(e - e_full) <= 1e-12 * abs(e_full) + 1e-18


def test_gemm_form_matches_loop_form(oracle):
    oracle.blas_init()
    for no, nv in [(4, 7), (6, 9), (10, 50)]:
        t = oracle.synthetic(no, nv, seed=5)
        e1, _, p1, _ = oracle.run(no, nv, t, per_tuple=True)
        e2, _, p2, _ = oracle.run(no, nv, t, use_gemm=True, per_tuple=True)
        assert np.max(np.abs(p1 - p2)) <= 1e-12 * np.max(np.abs(p1))
        assert abs(e1 - e2) <= 1e-12 * abs(e1)


@pytest.mark.parametrize("mistake", [m for m in
                                     ["p3_perm", "h0_sign", "drop_p4", "same_as_distinct",
                                      "factor", "h1_transpose", "singles_index"]])
def test_planted_mistake_breaks_identity(oracle, mistake):
    no, nv = 4, 7
    t = oracle.synthetic(no, nv, seed=1)
    e_full, _ = _full_space_energy(no, nv, t)
    good, _ = oracle.run(no, nv, t)
    bad, _ = oracle.run(no, nv, t, mistake=mistake)
    assert abs(good - e_full) <= 1e-12 * abs(e_full)
    assert abs(bad - e_full) > 1e-5 * abs(e_full)


def _slices(t, no, nv, a, b, c, J=False):
    """Slices of tuple (a,b,c) as src/atrip/Unions.hpp cuts them."""
    T = t["Tabij"].reshape((nv, nv, no, no), order="F")
    Vp = t["Jabci" if J else "Vabci"].reshape((nv, nv, nv, no), order="F")
    Vh = t["Jijka" if J else "Vijka"].reshape((no, no, no, nv), order="F")
    f = lambda x: np.asarray(x).reshape(-1, order="F")
    s = {}
    for name, (x, y) in {"VAB": (a, b), "VAC": (a, c), "VBC": (b, c), "VBA": (b, a),
                         "VCA": (c, a), "VCB": (c, b)}.items():
        s[name] = f(Vp[x, y, :, :])
    for name, x in {"A": a, "B": b, "C": c}.items():
        s["Vhhh" + name] = f(Vh[:, :, :, x])
        s["T" + name] = f(T[x, :, :, :])
    for name, (x, y) in {"TAB": (a, b), "TAC": (a, c), "TBC": (b, c)}.items():
        s[name] = f(T[x, y, :, :])
    return s


def test_W_simultaneous_permutation_symmetry(oracle):
    no, nv = 5, 9
    t = oracle.synthetic(no, nv, seed=11)
    a, b, c = 1, 4, 7
    W = lambda x, y, z: oracle.doubles(no, nv, _slices(t, no, nv, x, y, z)).reshape(
        (no, no, no), order="F")
    w = W(a, b, c)
    assert np.allclose(W(b, a, c), w.transpose(1, 0, 2), rtol=0, atol=1e-14)
    assert np.allclose(W(a, c, b), w.transpose(0, 2, 1), rtol=0, atol=1e-14)
    assert np.allclose(W(c, b, a), w.transpose(2, 1, 0), rtol=0, atol=1e-14)
    # the same W through the GEMM form
    wg = oracle.doubles(no, nv, _slices(t, no, nv, a, b, c), gemm=True)
    assert np.allclose(wg.reshape((no, no, no), order="F"), w, rtol=0, atol=1e-14)
    # and through the unrestricted numpy definition
    _, Wfull = _full_space_energy(no, nv, t)
    assert np.allclose(Wfull[a, b, c], w, rtol=0, atol=1e-14)


def test_distinct_energy_is_unrestricted_sum(oracle):
    no, nv = 6, 9
    rng = np.random.default_rng(0)
    w = rng.normal(size=no ** 3)
    z = w + 0.1 * rng.normal(size=no ** 3)
    epsi = rng.uniform(-1.5, -0.3, size=no)
    epsabc = 4.0
    e = oracle.energy(epsabc, no, epsi, w, z)
    W = w.reshape((no,) * 3, order="F")
    Z = z.reshape((no,) * 3, order="F")
    Y = (4 * Z + Z.transpose(1, 2, 0) + Z.transpose(2, 0, 1) - 2 * Z.transpose(0, 2, 1)
         - 2 * Z.transpose(1, 0, 2) - 2 * Z.transpose(2, 1, 0))
    D = epsabc - (epsi[:, None, None] + epsi[None, :, None] + epsi[None, None, :])
    assert abs(e - np.sum(2 * W * Y / D)) <= 1e-12 * abs(e)


def test_invariants(oracle):
    no, nv = 4, 7
    t = oracle.synthetic(no, nv, seed=2)
    e0, _ = oracle.run(no, nv, t)
    # uniform shift of all orbital energies leaves every denominator unchanged
    sh = dict(t, ei=t["ei"] + 0.37, ea=t["ea"] + 0.37)
    assert abs(oracle.run(no, nv, sh)[0] - e0) <= 1e-12 * abs(e0)
    # V = 0 gives E = 0
    zero = dict(t, Vabci=0 * t["Vabci"], Vijka=0 * t["Vijka"], Vabij=0 * t["Vabij"])
    assert oracle.run(no, nv, zero)[0] == 0.0
    # E(l T2, m V) = l^2 m^2 E(T2, V) at Tai = 0
    base = dict(t, Tai=0 * t["Tai"])
    eb, _ = oracle.run(no, nv, base)
    lam, mu = 1.7, 0.6
    sc = dict(base, Tabij=lam * t["Tabij"], Vabci=mu * t["Vabci"], Vijka=mu * t["Vijka"],
              Vabij=mu * t["Vabij"])
    assert abs(oracle.run(no, nv, sc)[0] - lam ** 2 * mu ** 2 * eb) <= 1e-12 * abs(eb)


def test_ijkabc_and_ranges(oracle):
    no, nv = 3, 8
    t = oracle.synthetic(no, nv, seed=9)
    n = oracle.n_tuples(nv)
    e, _, pt, _ = oracle.run(no, nv, t, per_tuple=True)
    # ijkabc: Tai negated, energy not negated (src/atrip/Atrip.cxx:183-187, :1108)
    neg = dict(t, Tai=-t["Tai"])
    e_neg, _ = oracle.run(no, nv, neg)
    e_ijk, _ = oracle.run(no, nv, t, ijkabc=True)
    assert e_ijk == -e_neg
    # sub-ranges add up and match the per-tuple energies
    for first, count in [(0, 5), (5, n - 5), (17, 40)]:
        es, _ = oracle.run(no, nv, t, first=first, count=count)
        assert abs(es + pt[first:first + count].sum()) <= 1e-13 * abs(e)


def test_cT_with_J_equal_V(oracle):
    no, nv = 3, 6
    t = oracle.synthetic(no, nv, seed=4, with_j=True)
    e, ect = oracle.run(no, nv, t)
    same = dict(t, Jabci=t["Vabci"], Jijka=t["Vijka"])
    e2, ect2 = oracle.run(no, nv, same)
    assert e2 == e and ect2 == e
    assert ect != e
    # linear in J: E_ct(2 J) = 2 E_ct(J) - ... bilinear in (W_J, Z): scales by 2
    dbl = dict(t, Jabci=2 * t["Jabci"], Jijka=2 * t["Jijka"])
    assert abs(oracle.run(no, nv, dbl)[1] - 2 * ect) <= 1e-12 * abs(ect)


# ---- complex (T) ---------------------------------------------------------

def _full_space_energy_z(no, nv, t, conj_holes=True, conj_w=True, real_ea=True):
    """Complex analogue of _full_space_energy: hole terms with conj(Vhhh),
    energy -(1/6) Re sum 2 conj(W) Y / D, with D = Re(ea_a + ea_b + ea_c)
    - (ei_i + ei_j + ei_k) (the reference takes the real part of the virtual
    sum only).  The switches drop either conjugation or keep Im(ea) (used to
    show the identity tells them apart)."""
    T = t["Tabij"].reshape((nv, nv, no, no), order="F")
    Vab = t["Vabij"].reshape((nv, nv, no, no), order="F")
    Vh = t["Vijka"].reshape((no, no, no, nv), order="F")
    Vh = np.conj(Vh) if conj_holes else Vh
    Vp = t["Vabci"].reshape((nv, nv, nv, no), order="F")
    tai = t["Tai"].reshape((nv, no), order="F")
    ei, ea = t["ei"], t["ea"]
    es = np.einsum
    W = (es("aEij,bcEk->abcijk", T, Vp) + es("aEik,cbEj->abcijk", T, Vp)
         + es("cEki,abEj->abcijk", T, Vp) + es("cEkj,baEi->abcijk", T, Vp)
         + es("bEji,acEk->abcijk", T, Vp) + es("bEjk,caEi->abcijk", T, Vp)
         - es("abLj,ikLc->abcijk", T, Vh) - es("abiL,jkLc->abcijk", T, Vh)
         - es("acLk,ijLb->abcijk", T, Vh) - es("aciL,kjLb->abcijk", T, Vh)
         - es("bcLk,jiLa->abcijk", T, Vh) - es("bcjL,kiLa->abcijk", T, Vh))
    S = (es("ai,bcjk->abcijk", tai, Vab) + es("bj,acik->abcijk", tai, Vab)
         + es("ck,abij->abcijk", tai, Vab))
    Z = W + S
    tr = lambda X, p: X.transpose(0, 1, 2, *[3 + q for q in p])
    Y = (4 * Z + tr(Z, (1, 2, 0)) + tr(Z, (2, 0, 1)) - 2 * tr(Z, (0, 2, 1))
         - 2 * tr(Z, (1, 0, 2)) - 2 * tr(Z, (2, 1, 0)))
    ev = np.real(ea) if real_ea else ea
    D = (ev[:, None, None, None, None, None] + ev[None, :, None, None, None, None]
         + ev[None, None, :, None, None, None] - ei[None, None, None, :, None, None]
         - ei[None, None, None, None, :, None] - ei[None, None, None, None, None, :])
    Wc = np.conj(W) if conj_w else W
    return -np.real(np.sum(2 * Wc * Y / D)) / 6.0


@pytest.mark.parametrize("no,nv", SIZES)
def test_complex_full_space_identity(oracle, no, nv):
    t = oracle.synthetic_z(no, nv, seed=3)
    e, ect = oracle.run_z(no, nv, t)
    assert ect == e
    e_full = _full_space_energy_z(no, nv, t)
    assert abs(e - e_full) <= 1e-12 * abs(e_full) + 1e-18


def test_complex_identity_sees_each_conjugation(oracle):
    no, nv = 4, 7
    t = oracle.synthetic_z(no, nv, seed=3)
    e, _ = oracle.run_z(no, nv, t)
    for kw in ({"conj_holes": False}, {"conj_w": False}):
        assert abs(e - _full_space_energy_z(no, nv, t, **kw)) > 1e-5 * abs(e), kw


def test_complex_epsabc_is_real_part(oracle):
    """Orbital energies with imaginary parts: the identity holds with the
    real part of ea_a + ea_b + ea_c and complex ei in the denominator
    (src/atrip/Atrip.cxx:643-644), and keeping Im(ea) would fail it."""
    no, nv = 4, 7
    t = oracle.synthetic_z(no, nv, seed=13)
    rng = np.random.default_rng(5)
    t["ea"] = t["ea"] + 1j * rng.uniform(0.05, 0.3, nv)
    t["ei"] = t["ei"] + 1j * rng.uniform(-0.2, 0.2, no)
    e, _ = oracle.run_z(no, nv, t)
    assert abs(e - _full_space_energy_z(no, nv, t)) <= 1e-12 * abs(e)
    assert abs(e - _full_space_energy_z(no, nv, t, real_ea=False)) > 1e-6 * abs(e)


def test_complex_with_real_inputs_is_the_real_path(oracle):
    for no, nv in [(3, 5), (5, 9)]:
        t = oracle.synthetic(no, nv, seed=8, with_j=True)
        tz = {k: v.astype(np.complex128) for k, v in t.items()}
        e, ect, p, pc = oracle.run(no, nv, t, per_tuple=True)
        ez, ectz, pz, pcz = oracle.run_z(no, nv, tz, per_tuple=True)
        assert np.max(np.abs(pz - p)) <= 1e-13 * np.max(np.abs(p))
        assert abs(ez - e) <= 1e-13 * abs(e) and abs(ectz - ect) <= 1e-13 * abs(ect)


def test_complex_generator_parts(oracle):
    t = oracle.synthetic(5, 8, seed=4, with_j=True)
    tz = oracle.synthetic_z(5, 8, seed=4, with_j=True)
    for k in t:
        assert np.array_equal(tz[k].real, t[k])
    assert not np.any(tz["ei"].imag) and not np.any(tz["ea"].imag)
    for k in ("Tai", "Tabij", "Vabij", "Vijka", "Vabci", "Jijka", "Jabci"):
        assert np.std(tz[k].imag) > 0.5 * np.std(tz[k].real)
    T = tz["Tabij"].reshape((8, 8, 5, 5), order="F")
    assert np.array_equal(T, T.transpose(1, 0, 3, 2))  # Tabij = Tbaji


def test_complex_cT_and_ijkabc(oracle):
    no, nv = 4, 6
    t = oracle.synthetic_z(no, nv, seed=9, with_j=True)
    tJ = dict(t)
    tJ["Jijka"], tJ["Jabci"] = t["Vijka"], t["Vabci"]
    e, ect = oracle.run_z(no, nv, tJ)
    assert abs(ect - e) <= 1e-14 * abs(e)  # J = V gives E(cT) = E(T)
    e1, _ = oracle.run_z(no, nv, t, ijkabc=True)
    tn = dict(t)
    tn["Tai"] = -t["Tai"]
    e2, _ = oracle.run_z(no, nv, tn)
    assert abs(e1 + e2) <= 1e-14 * abs(e1)  # ijkabc: Tai negated, no final sign


# ---- the slice-wise synthetic loop (large shapes) -------------------------

@pytest.mark.parametrize("no,nv,seed,with_j,scale", [
    (5, 11, 3, False, 1.0), (6, 13, 4, True, 1.0), (7, 9, 5, True, 0.37),
    (4, 20, 6, False, 1.0)])
def test_slice_wise_synthetic_loop_is_the_whole_tensor_loop(oracle, no, nv, seed,
                                                            with_j, scale):
    """oracle_run_synthetic draws each tuple's slices from the generator;
    on whole synthetic tensors (amplitudes and integrals times `scale`) the
    whole-tensor loop gives the same per-tuple bits, in both forms and on a
    range that starts mid-space."""
    t = oracle.synthetic(no, nv, seed, with_j=with_j)
    for k in ("Tai", "Tabij", "Vabij", "Vijka", "Vabci", "Jijka", "Jabci"):
        if k in t:
            t[k] = t[k] * scale
    n = oracle.n_tuples(nv)
    for gemm in (False, True):
        _, _, pt, ptc = oracle.run(no, nv, t, per_tuple=True, use_gemm=gemm)
        for first, count in [(0, n), (n // 3, 17)]:
            a, b = oracle.run_synthetic(no, nv, seed, first, count, with_j=with_j,
                                        scale=scale, use_gemm=gemm, nthreads=3)
            assert np.array_equal(a, pt[first:first + count])
            assert np.array_equal(b, ptc[first:first + count])
    # ijkabc negates Tai in both
    _, _, pt, _ = oracle.run(no, nv, t, per_tuple=True, ijkabc=True)
    a, _ = oracle.run_synthetic(no, nv, seed, 0, n, with_j=with_j, scale=scale,
                                ijkabc=True, use_gemm=False)
    assert np.array_equal(a, pt)
