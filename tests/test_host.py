"""Host-side logic: parameter tables, batch layout, C-ABI library surface."""

import ctypes as ct

import numpy as np
import torch
import pytest

import dxtb_amd._lib as L
from dxtb_amd import workloads as W
from dxtb_amd.engine import BatchLayout, alpha_beta
from dxtb_amd.param import SpeciesTables, slater_to_gauss

import oracle.gfn1 as og


def test_library_loads_and_exports_every_symbol():
    lib = L.load()
    for name in L.EXPORTED:
        assert hasattr(lib, name)
    # every function declared in include/xtb_hip.h is exported
    import re
    from pathlib import Path
    hdr = (Path(__file__).resolve().parent.parent / "include" / "xtb_hip.h").read_text()
    declared = set(re.findall(r"^int32_t (xtb_\w+)\(", hdr, flags=re.M))
    assert declared == set(L.EXPORTED)
    sizes = (ct.c_int64 * 4)()
    assert lib.xtb_abi_struct_sizes(sizes) == 0
    assert list(sizes) == [ct.sizeof(s) for s in (L.XtbBatch, L.XtbParams, L.XtbScfOpts, L.XtbScfState)]


def test_species_tables_match_oracle_basis():
    tb = SpeciesTables([1, 6, 7, 8])
    for z in (1, 6, 7, 8):
        bas = og.build_basis([z])
        s = tb.sp_of_z[z]
        base = tb.us_base[s]
        for k in range(tb.nsh_of_sp[s]):
            n = tb.us_nprim[base + k]
            a = tb.us_alpha[(base + k) * 8:(base + k) * 8 + n]
            c = tb.us_coeff[(base + k) * 8:(base + k) * 8 + n]
            assert np.allclose(a, bas.shell_alpha[k].numpy(), rtol=1e-15, atol=0)
            assert np.allclose(c, bas.shell_coeff[k].numpy(), rtol=1e-13, atol=0)
    K = og._shell_pair_k(og.build_basis([1, 6, 7, 8]), og.load_params()).numpy()
    assert np.allclose(tb.us_K, K, rtol=1e-15, atol=0)


def test_batch_layout_ragged_and_padded():
    mols = W.config1(count=5)
    zs = [z for z, _ in mols]
    tb = SpeciesTables(np.unique(np.concatenate(zs)))
    for padded in (False, True):
        lay = BatchLayout(zs, tb, padded=padded, chrg=np.zeros(5))
        a = lay.arrays
        for b, z in enumerate(zs):
            bas = og.build_basis(z)
            assert lay.nao[b] == bas.nao and lay.nsh[b] == bas.nsh
            v0 = lay.vec0[b]
            assert list(a["ao_sh"][v0:v0 + bas.nao]) == bas.ao_shell
            assert np.all(a["ao_sh"][v0 + bas.nao:lay.vec0[b + 1]] == -1)
            assert lay.n[b] == (lay.nao.max() if padded else bas.nao)
            s0 = lay.sh0[b]
            assert list(a["sh_atom"][s0:s0 + bas.nsh]) == bas.shell_atom
            nel = float(og._n0(bas).sum())
            assert a["mol_nel"][2 * b] + a["mol_nel"][2 * b + 1] == pytest.approx(nel)
        # tiles cover every upper-triangle shell pair exactly once
        t = a["tile_sh"].reshape(-1, 3)
        for b in range(5):
            m = lay.nsh[b]
            cov = np.zeros((m, m), dtype=int)
            for mm, i0, j0 in t[t[:, 0] == b]:
                for i in range(i0, min(i0 + 16, m)):
                    for j in range(j0, min(j0 + 16, m)):
                        if j >= i:
                            cov[i, j] += 1
            assert np.array_equal(np.triu(cov), np.triu(np.ones((m, m), dtype=int)))
        assert np.all(np.diff(lay.mat0) % 32 == 0)


def test_alpha_beta_split():
    na, nb = alpha_beta(np.array([10.0, 9.0, 9.0]), None)
    assert list(na) == [5.0, 5.0, 5.0] and list(nb) == [5.0, 4.0, 4.0]
    na, nb = alpha_beta(np.array([10.0]), 2)
    assert (na[0], nb[0]) == (6.0, 4.0)


def test_sto_expansion_identical_to_oracle():
    for (ng, n, l) in [(6, 2, 0), (6, 2, 1), (4, 1, 0), (3, 2, 0)]:
        a, c = slater_to_gauss(ng, n, l, 1.7)
        ao, co = og.slater_to_gauss(ng, n, l, 1.7)
        assert np.allclose(a, ao.numpy(), rtol=1e-15, atol=0) and np.allclose(c, co.numpy(), rtol=1e-15, atol=0)


def test_large_path_workspace_query():
    """xtb_scf_large_workspace (host-only): one molecule, the size of its
    scratch grows like n^2, and orbital counts beyond the sort's limit or
    batches of several molecules are refused before any device work."""
    lib = L.load()
    out = ct.c_int64(0)
    sizes = []
    for n in (700, 1400, 4800):
        b = L.XtbBatch(nmol=1, max_n=n, max_nsh=n // 2, max_nat=n // 3)
        assert lib.xtb_scf_large_workspace(ct.byref(b), ct.byref(out)) == L.XTB_OK
        assert out.value >= n * n + 4 * n
        sizes.append(out.value)
    assert 3.5 < sizes[1] / sizes[0] < 4.5
    b = L.XtbBatch(nmol=1, max_n=8193, max_nsh=4000, max_nat=3000)
    assert lib.xtb_scf_large_workspace(ct.byref(b), ct.byref(out)) == L.XTB_ETOOBIG
    b = L.XtbBatch(nmol=2, max_n=800, max_nsh=400, max_nat=300)
    assert lib.xtb_scf_large_workspace(ct.byref(b), ct.byref(out)) == L.XTB_EINVAL


def test_warm_start_flags():
    """A warm-start run keeps its state; it starts warm only when the
    pipeline holds a kept state from the previous run."""
    from dxtb_amd.engine import ScfOptions

    o = ScfOptions(warm_start=True)
    s = o.to_struct(want_w=True, padded=False, warm=False)
    assert (s.keep_state, s.warm_start) == (1, 0)
    s = o.to_struct(want_w=True, padded=False, warm=True)
    assert (s.keep_state, s.warm_start) == (1, 1)
    s = ScfOptions().to_struct(want_w=False, padded=False, warm=True)
    assert (s.keep_state, s.warm_start) == (0, 0)


def test_tile_lists_cover_each_pair_space_once():
    """The 16 x 16 tiles of every molecule's upper pair triangle, vectorised
    over the batch: each (i0 <= j0) block exactly once, in molecule order."""
    from dxtb_amd.engine import _tiles

    counts = np.array([0, 5, 17, 33, 1, 0, 100])
    t = _tiles(counts)
    want = []
    for m, c in enumerate(counts):
        k = (c + 15) // 16
        want += [(m, 16 * i, 16 * j) for i in range(k) for j in range(i, k)]
    assert [tuple(r) for r in t] == want
    assert _tiles(np.array([0, 0])).shape == (0, 3)


def test_pair_scaling_derivatives_match_finite_differences():
    """param.shell_pair_k returns K_ab and its derivatives in the Hamiltonian
    section's parameters and the electronegativities (the chain rule behind
    the calculator's parameter gradients); central differences agree."""
    import copy

    from dxtb_amd.param import SpeciesTables, elements_by_z, load_gfn1, shell_pair_k

    tb = SpeciesTables([1, 6, 7, 8])
    ham = load_gfn1()["hamiltonian"]
    en = {z: e["en"] for z, (_, e) in elements_by_z().items()}
    K, dks = shell_pair_k(tb, ham, en, grad=True)
    assert np.array_equal(K, tb.us_K)
    names = {n for n, _ in dks}
    assert {("shell", "ss"), ("shell", "sp"), ("shell", "pp"), ("kpol",), ("enscale",), ("kpair", "H-H"),
            ("en", 1), ("en", 8)} <= names
    h = 1e-6
    for name, dk in dks:
        ks = []
        for s in (1.0, -1.0):
            hh, ee = copy.deepcopy(ham), dict(en)
            if name[0] in ("shell", "kpair"):
                hh[name[0]][name[1]] += s * h
            elif name[0] == "en":
                ee[name[1]] += s * h
            else:
                hh[name[0]] += s * h
            ks.append(shell_pair_k(tb, hh, ee))
        assert np.abs((ks[0] - ks[1]) / (2 * h) - dk).max() < 1e-8, name


def test_scf_rejects_an_out_of_range_xcd_shift():
    """xtb_scf validates wg_skip (0..7) before it touches the device."""
    lib = L.load()
    b = L.XtbBatch(nmol=1, max_n=4)
    p = L.XtbParams()
    st = L.XtbScfState(**{f: 8 for f, _ in L.XtbScfState._fields_})
    from dxtb_amd.engine import ScfOptions

    o = ScfOptions().to_struct(want_w=False, padded=False, wg_skip=8)
    rc = lib.xtb_scf(ct.byref(b), ct.byref(p), 8, 8, 8, ct.byref(o), ct.byref(st), None)
    assert rc == 1  # invalid argument


def test_alpha_beta_rejects_spins_that_do_not_fit():
    """filling.py:80-93: more unpaired electrons than electrons, or odd/even
    parity mismatch, raise ValueError (they used to be replaced silently)."""
    with pytest.raises(ValueError):
        alpha_beta(np.array([8.0]), 1)
    with pytest.raises(ValueError):
        alpha_beta(np.array([3.0]), 5)
    with pytest.raises(ValueError):
        BatchLayout([np.array([8, 1, 1])], SpeciesTables([1, 8]), padded=False,
                    chrg=np.zeros(1), spin=np.array([1.0]))
    na, nb = alpha_beta(np.array([17.0, 8.0, 7.5]), None)
    assert na.tolist() == [9.0, 4.0, 3.75] and nb.tolist() == [8.0, 4.0, 3.75]
    na, nb = alpha_beta(np.array([8.0]), 2)
    assert (na[0], nb[0]) == (5.0, 3.0)


def test_negative_temperature_is_rejected():
    from dxtb_amd.engine import ScfOptions

    with pytest.raises(ValueError):
        ScfOptions(fermi_etemp=-1.0).to_struct(want_w=False, padded=False)
    ScfOptions(fermi_etemp=0.0).to_struct(want_w=False, padded=False)


def test_serving_merged_launch_takes_the_oldest_requests():
    """Saturation launches (merge_above): the oldest merge_cap waiting
    requests whatever their class, the larger class first among requests of
    one arrival time; a launch that does not fit the memory budget puts its
    requests back at the head of their classes in their order."""
    from collections import deque

    from dxtb_amd.serving import DEFAULT_EDGES, RaggedServer, Request

    srv = object.__new__(RaggedServer)
    srv.edges = DEFAULT_EDGES
    srv.queues = [deque() for _ in DEFAULT_EDGES]
    srv.large_queue = deque()
    srv.merge_cap = 5
    srv.merge_largest = False
    z = np.array([1, 1])
    mk = lambda rid, nao, t: Request(rid, z, np.zeros((2, 3)), arrival=t, nao=nao)
    first = [mk(0, 10, 0.0), mk(1, 20, 0.0), mk(2, 30, 0.0), mk(3, 150, 0.0), mk(4, 140, 0.0),
             mk(5, 700, 0.0)]
    later = [mk(6 + i, 60, 1.0) for i in range(4)]
    for r in first + later:
        srv.submit(r)
    assert srv.pending() == 10
    # above the largest class: the whole-GPU lane's queue, never a merged launch
    srv.submit(mk(99, DEFAULT_EDGES[-1] + 1, 0.0))
    assert srv.pending() == 11 and [r.rid for r in srv.large_queue] == [99]
    srv.large_queue.clear()
    take = srv._take_merged()
    # class 5 (700), class 2 (150, 140), then class 0 in arrival order
    assert [r.rid for r in take] == [5, 3, 4, 0, 1]
    for r in reversed(take):
        srv.submit_front(r)
    assert [r.rid for r in srv.queues[0]] == [0, 1, 2]
    assert [r.rid for r in srv.queues[2]] == [3, 4]
    take = srv._take_merged()
    assert [r.rid for r in take] == [5, 3, 4, 0, 1]
    take = srv._take_merged()
    assert [r.rid for r in take] == [2, 6, 7, 8, 9]
    assert srv.pending() == 0 and srv._take_merged() == []
    # largest first: class 5, class 2, class 1 (later arrivals), class 0
    srv.merge_largest = True
    for r in first + later:
        srv.submit(r)
    assert [r.rid for r in srv._take_merged()] == [5, 3, 4, 6, 7]
    assert [r.rid for r in srv._take_merged()] == [8, 9, 0, 1, 2]
