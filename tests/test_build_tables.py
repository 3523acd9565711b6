"""Host logic of the build: ket-tile rule and the per-class tables it reads."""

import json
import os
import re

import pytest

from _util import PKG  # noqa: F401  (puts the package on sys.path)
from vlxmi.gen import build as B
from vlxmi.gen import emit
from vlxmi.pairs import j_classes, k_classes


def test_tile_rule_reaches_register_occupancy():
    # 512 registers -> 1 wave/SIMD -> 4 waves/CU -> LDS budget 40 KB
    assert B.tile_for(512, 41216, 32, 712) == 16          # 29.8 KB after halving once
    assert B.tile_for(512, 59648, 32, 712) == 4           # fixed part 36.9 KB leaves room for 4 records
    # 252 registers -> 2 waves/SIMD -> 8 waves/CU -> 20 KB
    assert B.tile_for(252, 44032, 64, 688) == 16
    # 68 registers -> 7 waves/SIMD -> 28 waves/CU -> 5851 B
    assert B.tile_for(68, 8192, 64, 128) == 32
    # already within budget: unchanged
    assert B.tile_for(512, 30000, 32, 712) == 32
    # never below the minimum tile
    assert B.tile_for(512, 100000, 32, 712) == 4


def _heavy(kind, cls):
    la, lb, lc, ld = [emit.LV[x] for x in cls]
    return len(emit.plan_slices(la, lb, lc, ld, "J" if kind == "J" else "Kheavy")) > 1


@pytest.mark.parametrize("path,kind", [(B.J_LAYOUT_TABLE, "J"), (B.LAYOUT_TABLE, "K")])
def test_layout_tables_name_split_classes(path, kind):
    lay = json.load(open(path))
    classes = [ab + cd for ab, cd in j_classes()] if kind == "J" else list(k_classes())
    for cls, v in lay.items():
        assert cls in classes and v in ("reg", "lds")
        assert _heavy(kind, cls), cls          # only classes that run in passes have two layouts


def test_tile_table_entries():
    tiles = json.load(open(B.TILE_TABLE))
    names = set(B.sources(tiles={}, layouts={}))
    for fn, kt in tiles.items():
        assert fn in names
        assert kt in (4, 8, 16, 32), (fn, kt)


def test_sources_use_table_values():
    """Kernels take their ket tile from the tile table, except those that keep
    the coarse Boys rows in LDS (a fixed 16-record tile)."""
    tiles = B.load_tiles()
    blds = B.load_boys_lds()
    srcs = B.sources()
    checked = 0
    for fn, kt in tiles.items():
        pre = fn.split("_")[0]
        fam = ("J" if pre.startswith("j") else "K") + ("b" if "b" in pre[1:] else "")
        cls = fn.split("_")[1].split(".")[0]
        m = re.search(r"constexpr int KT = (\d+)", srcs[fn])
        want = emit.LDS_BOYS_KT if (cls in blds[fam] and not fn.startswith("schwarz")) else kt
        assert m and int(m.group(1)) == want, fn
        checked += 1
    assert checked == len(tiles)
    for fam, pres in (("J", ("j", "jm")), ("K", ("k", "km")), ("Jb", ("jb", "jbm")), ("Kb", ("kb", "kbm"))):
        for cls in blds[fam]:
            for pre in pres:
                s = srcs["%s_%s.hip" % (pre, cls)]
                assert "vlxmi::boys_lds<" in s and "vlxmi::boys<" not in s and "__syncthreads();" in s


def test_waves_per_workgroup_table():
    """Classes run with several wavefronts per workgroup only where the table
    says so; those kernels synchronise each wavefront on its own after the one
    workgroup barrier that publishes the LDS Boys rows, and are launched with
    64 * waves threads."""
    wp = B.load_wpg()
    srcs = B.sources()
    for fam, pres in (("J", ("j", "jm")), ("K", ("k", "km")), ("Jb", ("jb", "jbm")), ("Kb", ("kb", "kbm"))):
        for pre in pres:
            for fn, s in srcs.items():
                if not fn.startswith(pre + "_"):
                    continue
                cls = fn.split("_")[1].split(".")[0]
                w = int(wp[fam].get(cls, 1))
                if w == 1:
                    assert "vlxmi::wave_sync()" not in s and "WPG" not in s, fn
                else:
                    assert "constexpr int WPG = %d;" % w in s and "vlxmi::wave_sync();" in s, fn
                    assert s.count("__syncthreads();") <= 1, fn
                    assert "dim3(%d)" % (64 * w) in s, fn


def test_mixed_kernels_have_one_body_per_loop():
    """The FP64 loop switches to the FP32 loop once; no per-record branch
    between two bodies remains."""
    for cls in ("ssss", "pdpd", "ddsp"):
        src = emit.emit_j(cls, mixed=True)
        assert src.count("if (__all(!active || qmax * R[0] < fp32_threshold)) { lo_mode = true; break; }") == 1
        assert "if (__all(lo))" not in src
    src = emit.emit_k("pddd", mixed=True)
    npass = len(emit.plan_slices(1, 2, 2, 2, "Kheavy"))
    assert src.count("< fp32_threshold)) { lo") == npass


def test_density_order_table():
    """Classes in k_dorder.json read D[l][j] (one density row per wave); all
    others read D[j][l] as the reference does."""
    tab = B.load_layouts(B.DORDER_TABLE)
    assert tab and set(tab.values()) == {"lj"}
    kc = list(k_classes())
    srcs = B.sources()
    for cls in kc:
        for pre in ("k", "km", "kb", "kbm"):
            s = srcs["%s_%s.hip" % (pre, cls)]
            lj = "Dm[(size_t)ket_cg[s * ND + 0] * nao + " in s
            jl = re.search(r"Dm\[\(size_t\)\w+(\[\d+\])? \* nao \+ ket_cg\[s \* ND \+ 0\]\]", s) is not None
            assert (lj, jl) == ((True, False) if cls in tab else (False, True)), (pre, cls)
    for cls in tab:
        assert cls in kc


def test_mixed_policy_table():
    from vlxmi import policy
    tab = policy.mixed_classes()
    assert set(tab) == {"J", "K"} and tab["J"] and tab["K"]
    assert tab["J"] <= {ab + cd for ab, cd in j_classes()}
    assert tab["K"] <= set(k_classes())
    cj, ck = sorted(tab["J"])[0], sorted(tab["K"])[0]
    assert policy.use_mixed("J", cj, "mixed") and policy.use_mixed("K", ck, "mixed")
    assert not policy.use_mixed("J", cj, "fp64")
    assert not policy.use_mixed("J", "ssss", "mixed") or "ssss" in tab["J"]
    assert policy.use_mixed("K", "ssss", "mixed", "all")
    with pytest.raises(ValueError):
        policy.use_mixed("K", "ssss", "mixed", "some")


def test_mixed_fp32_passes():
    """The FP32 part of a mixed kernel runs its own (never more) passes:
    FP32 intermediates take half the registers of FP64 ones.  A heavy exchange
    class whose FP32 recursion needs as many passes keeps one FP64-then-FP32
    loop pair per pass; one that needs fewer runs its FP32 passes after all
    FP64 passes."""
    for cls in ("sddd", "ppdp", "pddd"):
        la, lb, lc, ld = [emit.LV[x] for x in cls]
        n64 = len(emit.plan_slices(la, lb, lc, ld, "Kheavy"))
        n32 = len(emit.plan_slices_f32(la, lb, lc, ld, "Kheavy"))
        assert 1 <= n32 <= n64
        src = emit.emit_k(cls, mixed=True)
        assert src.count("< fp32_threshold)) { lo") == n64
        assert src.count("bool actf") == (n32 if n32 < n64 else 0)
    assert len(emit.plan_slices_f32(0, 2, 2, 2, "Kheavy")) < len(emit.plan_slices(0, 2, 2, 2, "Kheavy"))
    for cls in ("sddd", "pppd"):
        la, lb, lc, ld = [emit.LV[x] for x in cls]
        assert len(emit.plan_slices_f32(la, lb, lc, ld, "J")) < len(emit.plan_slices(la, lb, lc, ld, "J"))


def test_gradient_layout_table():
    """Every gradient class named in grad_passes.json has a known layout, and
    the generated sources follow it: Coulomb classes listed run one pass per
    ket-first-index component or per ket component pair; exchange classes
    split by centre have an _A and a _C kernel, launched one after the other
    by the class launcher, each with one pass, one per ket-first-index
    component or one per ket component pair."""
    from vlxmi.gen import emit_grad
    gp = B.load_grad_passes()
    assert set(gp["J"].values()) <= {"component", "pair"} and gp["J"]
    assert set(gp["K"].values()) <= {"component", "split_one", "split_component", "split_pair"} and gp["K"]
    assert set(gp["J"]) <= {ab + cd for ab, cd in j_classes()} and set(gp["K"]) <= set(k_classes())
    srcs = B.sources()
    for cls, v in gp["K"].items():
        passes, split = B.grad_k_layout(gp, cls)
        s = srcs["kg_%s.hip" % cls]
        name = "computeExchangeGradient" + cls.upper()
        if split:
            assert (name + "_A(") in s and (name + "_C(") in s
            assert s.count("hipLaunchKernelGGL(%s_" % name) == 2
        nc, nd = [(emit.LV[x] + 1) * (emit.LV[x] + 2) // 2 for x in cls[2:]]
        assert len(emit_grad.grad_groups(*[emit.LV[x] for x in cls], passes)) == \
            {"component": nc, "pair": nc * nd, "one": 1}[passes]
    for cls, v in gp["J"].items():
        assert "computeCoulombGradient" + cls.upper() in srcs["jg_%s.hip" % cls]
        nc, nd = [(emit.LV[x] + 1) * (emit.LV[x] + 2) // 2 for x in cls[2:]]
        assert len(emit_grad.grad_groups(*[emit.LV[x] for x in cls], v)) == (nc if v == "component" else nc * nd)


# Scratch per lane a kernel may use: the runtime backs it for every wave the
# device can hold, and a 34 KB-per-lane kernel failed to dispatch
# (out of resources) in a process with several queues.
SCRATCH_LIMIT = 16 * 1024


def test_kernel_scratch_within_limit():
    """Every compiled J/K and gradient kernel needs at most SCRATCH_LIMIT bytes
    of scratch per lane (register spills)."""
    import glob
    objs = glob.glob(os.path.join(B.OBJ, "*.hip.*.o"))
    if not objs:
        pytest.skip("library not built")
    worst = {}
    for o in objs:
        if not os.path.basename(o).startswith(("j_", "k_", "jm_", "km_", "jb", "kb", "jg_", "kg_")):
            continue
        for name, scr in B.kernel_scratch(o).items():
            worst[name] = scr
    assert worst
    over = {n: s for n, s in worst.items() if s > SCRATCH_LIMIT}
    assert not over, over
