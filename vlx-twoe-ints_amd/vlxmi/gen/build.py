"""Generate the HIP sources, the C header, and build libvlxmi.so for gfx950.

Usage: ``python -m vlxmi.gen.build [--jobs N] [--force] [--only CLASS...] [--retune]``

Outputs:
  vlx-twoe-ints_amd/csrc/generated/*.hip   emitted kernels + launchers
  include/vlxmi_kernels.h                  C declarations of every launcher
  vlx-twoe-ints_amd/vlxmi/_lib/libvlxmi.so  the shared library (in-tree)

Objects are cached by a hash of (source text, flags) under
vlx-twoe-ints_amd/_build/; ``force=True`` recompiles everything.

Ket-tile sizes.  A kernel's waves per CU are bounded by its register count
(512 per lane per SIMD, 4 SIMDs) and by LDS (160 KB per CU, one wavefront per
workgroup).  The LDS of a J/K kernel is dominated by its staged ket tile, so
``--retune`` reads each compiled kernel's register count and LDS size from
the code object, shrinks the ket tile (powers of two, at least 4 records) of
every kernel whose LDS would allow fewer waves than its registers do, and
records the sizes in ket_tiles.json beside this file, which later builds read.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import json
import os
import re
import shutil
import tempfile
import subprocess
import sys
import time

from . import emit
from . import emit_grad
from ..pairs import PAIR_CLASSES, j_classes, k_classes

PKG = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))   # vlx-twoe-ints_amd
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
GEN = os.path.join(CSRC, "generated")
OBJ = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "vlxmi", "_lib")
LIB = os.path.join(LIBDIR, "libvlxmi.so")
HEADER = os.path.join(ROOT, "include", "vlxmi_kernels.h")
DEVICE_HEADER = os.path.join(ROOT, "include", "vlxmi_device_kernels.h")

TILE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ket_tiles.json")
# heavy exchange classes: register layout per class ("reg" or "lds", see
# emit.emit_k), chosen from per-class timings of both layouts on lattice64
LAYOUT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "k_layouts.json")
# split Coulomb classes: accumulator/threshold layout per class ("reg" or "lds",
# see emit.emit_j), chosen the same way
J_LAYOUT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "j_layouts.json")
# exchange classes that read the density as D[l][j] ("lj", see emit.emit_k),
# chosen from per-class timings of both orders on lattice64
DORDER_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "k_dorder.json")
# light classes that evaluate the Boys function from coarse rows kept in LDS
# (see emit.lds_boys_setup), chosen from per-class timings on lattice64
BOYS_LDS_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "boys_lds.json")


# light classes whose kernels run several wavefronts per workgroup (sharing the
# LDS Boys rows), chosen from per-class timings
WPG_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wpg.json")

# heavy exchange classes whose plain FP64 kernel runs one pass per ket
# component pair (emit.emit_k kpairs), chosen from lattice64 per-class timings
K_PASSES_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "k_passes.json")


J_PASSES_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "j_passes.json")


def _pair_classes(path):
    if not os.path.exists(path):
        return set()
    with open(path) as f:
        return {k for k, v in json.load(f).items() if not k.startswith("_") and v == "pair"}


def load_k_passes():
    """Classes whose plain FP64 exchange kernel runs pair passes."""
    return _pair_classes(K_PASSES_TABLE)


def load_j_passes():
    """Split Coulomb classes whose plain FP64 kernel runs pair passes."""
    return _pair_classes(J_PASSES_TABLE)


# gradient classes whose kernels run one pass per ket-first-index component
# (emit_grad.grad_groups); the others run one pass.  Chosen from per-class
# timings of both on lattice64.
GRAD_PASSES_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "grad_passes.json")


def grad_k_layout(gp, cls):
    """(passes, split) of the exchange-gradient kernel of ``cls``."""
    v = gp["K"].get(cls, "one")
    return (v[len("split_"):], True) if v.startswith("split_") else (v, False)


def load_grad_passes():
    """{"J": {class: passes}, "K": {class: passes}}; absent classes run one pass."""
    tab = {}
    if os.path.exists(GRAD_PASSES_TABLE):
        with open(GRAD_PASSES_TABLE) as f:
            tab = json.load(f)
    return {k: dict(tab.get(k, {})) for k in ("J", "K")}


FAMILIES = ("J", "K", "Jb", "Kb")     # class launchers (+ _mixed), batched (+ _mixed)


def load_wpg():
    """{family: {class: waves per workgroup}} (families J, K, Jb, Kb); classes
    absent run one wavefront per workgroup."""
    tab = {}
    if os.path.exists(WPG_TABLE):
        with open(WPG_TABLE) as f:
            tab = json.load(f)
    return {fam: dict(tab.get(fam, {})) for fam in FAMILIES}


def load_boys_lds():
    """{family: set of classes} using the LDS Boys rows (families J, K, Jb, Kb)."""
    tab = {}
    if os.path.exists(BOYS_LDS_TABLE):
        with open(BOYS_LDS_TABLE) as f:
            tab = json.load(f)
    return {fam: set(tab.get(fam, [])) for fam in FAMILIES}
LLVM = "/opt/rocm/llvm/bin"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", CSRC, "-I", os.path.join(ROOT, "include")]


def j_name(ab, cd):
    return ab + cd


def load_tiles():
    if os.path.exists(TILE_TABLE):
        with open(TILE_TABLE) as f:
            return {k: int(v) for k, v in json.load(f).items()}
    return {}


def load_layouts(path=None):
    path = path or LAYOUT_TABLE
    if os.path.exists(path):
        with open(path) as f:
            return {k: v for k, v in json.load(f).items() if not k.startswith("_")}
    return {}


def sources(tiles=None, layouts=None):
    """Map file name -> source text of every generated translation unit."""
    tiles = load_tiles() if tiles is None else tiles
    layouts = load_layouts() if layouts is None else layouts
    jlayouts = load_layouts(J_LAYOUT_TABLE)
    dorders = load_layouts(DORDER_TABLE)
    blds = load_boys_lds()
    wpg = load_wpg()
    out = {}
    inc = '#include "vlxmi_device.hpp"\n\n'
    jpairs = load_j_passes()
    for ab, cd in j_classes():
        cls = j_name(ab, cd)
        for pre, kw in (("j", dict(jpairs=cls in jpairs)), ("jm", dict(mixed=True)), ("jb", dict(batched=True)),
                        ("jbm", dict(batched=True, mixed=True))):
            fn = "%s_%s.hip" % (pre, cls)
            fam = "Jb" if kw.get("batched") else "J"
            out[fn] = inc + emit.emit_j(cls, kt=tiles.get(fn), layout=jlayouts.get(cls, "reg"),
                                        lds_boys=cls in blds[fam], wpg=int(wpg[fam].get(cls, 1)), **kw)
    kpairs = load_k_passes()
    for cls in k_classes():
        for pre, kw in (("k", dict(kpairs=cls in kpairs)), ("km", dict(mixed=True)), ("kb", dict(batched=True)),
                        ("kbm", dict(batched=True, mixed=True))):
            fn = "%s_%s.hip" % (pre, cls)
            fam = "Kb" if kw.get("batched") else "K"
            out[fn] = inc + emit.emit_k(cls, kt=tiles.get(fn), layout=layouts.get(cls, "reg"),
                                        dorder=dorders.get(cls, "jl"), lds_boys=cls in blds[fam],
                                        wpg=int(wpg[fam].get(cls, 1)), **kw)
    gp = load_grad_passes()
    for ab, cd in j_classes():
        cls = j_name(ab, cd)
        out["jg_%s.hip" % cls] = inc + emit_grad.emit_jgrad(cls, passes=gp["J"].get(cls, "one"))
    for cls in k_classes():
        passes, split = grad_k_layout(gp, cls)
        out["kg_%s.hip" % cls] = inc + emit_grad.emit_kgrad(cls, passes=passes, split=split)
    for pc in PAIR_CLASSES:
        out["schwarz_%s.hip" % pc] = inc + emit.emit_schwarz(pc)
    out["jdispatch.hip"] = inc + jdispatch_text()
    out["splits.hip"] = inc + splits_text()
    return out


# VeloxChem launches the classes cut for register pressure as series of split
# kernels (computeCoulombFockPDDD0..6 etc.), all accumulating into the same
# output (summary_fock_split.txt:1-20; names made at split_coulomb_fock.py:242-249
# and new_exchange_code/exchangefocksplitter.py:167-176).
SPLIT_SERIES = {"J": {"pddd": 7, "ddpd": 10, "dddd": 30},
                "K": {"pddd": 8, "dpdd": 7, "dddp": 7, "dddd": 19}}


def splits_text():
    """Launchers named after the reference's split kernels: split 0 runs the
    whole class (these kernels are split into component passes inside one
    launch instead), splits 1..n-1 check their arguments and launch nothing,
    so a host that launches the whole series into one buffer gets the class's
    result."""
    o = []
    for kind, series in SPLIT_SERIES.items():
        for cls, n in series.items():
            if kind == "J":
                sig = emit._strip_restrict(emit.j_signature(cls))
                tail, targs = ["uint32_t ket_split", "void* stream"], ["ket_split", "stream"]
                base = "computeCoulombFock" + cls.upper()
                count = "%s_prim_pair_count_local" % cls[:2]
            else:
                sig = emit._strip_restrict(emit.k_signature(cls))
                tail, targs = ["void* stream"], ["stream"]
                base = "computeExchangeFock" + cls.upper()
                count = "pair_inds_count_for_K_%s" % (cls[0] + cls[2])
            args = [x.split()[-1] for x in sig]
            for k in range(n):
                o.append('extern "C" int vlxmi_launch_%s%d(%s)' % (base, k, ", ".join(sig + tail)))
                o.append("{")
                if k == 0:
                    o.append("    return vlxmi_launch_%s(%s);" % (base, ", ".join(args + targs)))
                else:
                    o.append("    (void)stream;")
                    for a in args:
                        o.append("    (void)%s;" % a)
                    if kind == "J":
                        o.append("    (void)ket_split;")
                    o.append("    (void)%s;" % count)
                    o.append("    return VLXMI_OK;")
                o.append("}")
                o.append("")
    return "#include \"vlxmi_kernels.h\"\n\n" + "\n".join(o) + "\n"


def device_header_text():
    """The kernels' own declarations (the reference prints its kernels'
    declarations with the ``header`` argument, gen_coulomb.py:38-45,95-98)."""
    o = ["/* Generated by vlxmi.gen.build -- declarations of the J/K kernels themselves,",
         " * with the reference's names and parameter lists (gen_coulomb.py:74-93,",
         " * gen_exchange.py:75-115), for HIP code that launches them directly.",
         " *",
         " * Launch geometry (these kernels' own, not the reference's 2-D tiles):",
         " *   computeCoulombFock{AB}{CD}: blockDim 64 (one wavefront), gridDim.x =",
         " *     ceil(bra blocks / 64), gridDim.y = number of ket-list chunks (with more",
         " *     than one the chunks add to mat_J atomically; the launchers in",
         " *     vlxmi_kernels.h add them in a fixed order instead);",
         " *   computeExchangeFock{A}{B}{C}{D}: blockDim 64, gridDim.x = (i,k) blocks.",
         " * The lists must be in the primitive-pair block order (vlxmi_*_to_blocks",
         " * convert reference-order lists); the launchers check arguments and are the",
         " * supported entry points. */",
         "#ifndef VLXMI_DEVICE_KERNELS_H",
         "#define VLXMI_DEVICE_KERNELS_H",
         "#include <hip/hip_runtime.h>",
         "#include <stdint.h>",
         ""]
    for ab, cd in j_classes():
        cls = j_name(ab, cd)
        o.append("__global__ void computeCoulombFock%s(%s);" % (cls.upper(), ", ".join(emit._strip_restrict(emit.j_signature(cls)))))
    for cls in k_classes():
        o.append("__global__ void computeExchangeFock%s(%s);" % (cls.upper(), ", ".join(emit._strip_restrict(emit.k_signature(cls)))))
    o += ["",
          "/* Energy-gradient kernels (no counterpart kernel in the reference, which carries the",
          " * derivative operator only symbolically, eri.py:85-175):",
          " *   computeCoulombGradient{AB}{CD}: blockDim 64, gridDim (ceil(bra blocks / 64), ket chunks);",
          " *     chunk y stores the six sums of bra block bb at grad[(y * nbra + bb) * 6 + 0..5];",
          " *   computeExchangeGradient{A}{B}{C}{D}: blockDim 64, gridDim = (i,k) blocks; adds six sums",
          " *     per block to grad[blk * 6 + 0..5]; for the classes split by centre, the _A kernel adds",
          " *     columns 0..2 and the _C kernel columns 3..5 (the launcher runs both). */"]
    for ab, cd in j_classes():
        cls = j_name(ab, cd)
        o.append("__global__ void computeCoulombGradient%s(%s);" % (cls.upper(), ", ".join(emit._strip_restrict(emit_grad.jg_signature(cls)))))
    gp = load_grad_passes()
    for cls in k_classes():
        sig = ", ".join(emit._strip_restrict(emit_grad.kg_signature(cls)))
        if grad_k_layout(gp, cls)[1]:
            o.append("__global__ void computeExchangeGradient%s_A(%s);" % (cls.upper(), sig))
            o.append("__global__ void computeExchangeGradient%s_C(%s);" % (cls.upper(), sig))
        else:
            o.append("__global__ void computeExchangeGradient%s(%s);" % (cls.upper(), sig))
    o += ["", "#endif", ""]
    return "\n".join(o)


def jdispatch_text():
    """vlxmi_jpart_dispatch: the plain Coulomb launch of a class named by its
    four-letter string, with the primitive tables passed per angular momentum
    (index 0 s, 1 p, 2 d) and a caller-provided partial-sum buffer."""
    common = ["const double* cd_mat_D", "const double* ab_mat_Q_local", "const double* cd_mat_Q",
              "const uint32_t* ab_first_inds_local", "const uint32_t* ab_second_inds_local",
              "const double* ab_pair_data_local", "uint32_t ab_prim_pair_count_local",
              "const uint32_t* cd_first_inds", "const uint32_t* cd_second_inds", "const double* cd_pair_data",
              "uint32_t cd_prim_pair_count", "const double* boys_func_table", "const double* boys_func_ft",
              "double eri_threshold"]
    o = ["#include <cstring>", ""]
    for ab, cd in j_classes():
        cls = j_name(ab, cd)
        sig = emit._strip_restrict(emit.j_signature(cls))
        o.append('extern "C" int vlxmi_jpart_computeCoulombFock%s(%s,\n    uint32_t ket_split, double* part, void* stream);'
                 % (cls.upper(), ", ".join(sig)))
    o.append("")
    o.append('extern "C" int vlxmi_jpart_dispatch(const char* cls, double* mat_J, const double* const* prim_info,')
    o.append("    const uint32_t* prim_count, %s,\n    uint32_t ket_split, double* part, void* stream)" % ", ".join(common))
    o.append("{")
    names = [c.split()[-1] for c in common]
    for ab, cd in j_classes():
        cls = j_name(ab, cd)
        args = ["mat_J"]
        for x in "spd":
            if x in cls:
                l = "spd".index(x)
                args += ["prim_info[%d]" % l, "prim_count[%d]" % l]
        args += names
        o.append('    if (std::strcmp(cls, "%s") == 0)' % cls)
        o.append("        return vlxmi_jpart_computeCoulombFock%s(%s, ket_split, part, stream);" % (cls.upper(), ", ".join(args)))
    o.append("    return VLXMI_ERR_ARG;")
    o.append("}")
    return "\n".join(o) + "\n"


def header_text():
    o = ["/* Generated by vlxmi.gen.build -- C ABI of the per-class J/K launchers.",
         " *",
         " * Each vlxmi_launch_computeCoulombFock{AB}{CD} takes exactly the argument list",
         " * of the reference kernel computeCoulombFock{AB}{CD} (gen_coulomb.py:74-93) and",
         " * each vlxmi_launch_computeExchangeFock{A}{B}{C}{D} that of",
         " * computeExchangeFock{A}{B}{C}{D} (gen_exchange.py:75-115,",
         " * new_exchange_code/exchangefockwriter.py:55-104), all pointers device",
         " * pointers owned by the caller, outputs accumulated (+=), followed by launch",
         " * controls (J: ket_split = number of ket-list chunks run in parallel; with >1",
         " * each chunk's partial sums go to a workspace the library keeps per (device,",
         " * stream) and are added to mat_J in chunk order by a compensated reduction,",
         " * so results are bit-reproducible) and the hipStream_t as void*.  They replace the",
         " * kernel launches of VeloxChem's GPU Fock driver (src/gpu/EriCoulomb.cu /",
         " * EriExchange.cu, named at new_exchange_code/erisplitcounter.py:44).",
         " * Return: 0 ok, <0 argument/layout error (vlxmi_status.h), >0 hipError_t.",
         " * Lists must follow the primitive-pair block order documented in",
         " * vlx-twoe-ints_amd/vlxmi/pairs.py.",
         " * mat_D_full_AO must be symmetric (as every SCF density is): the exchange",
         " * classes listed in vlxmi/gen/k_dorder.json read D_jl as D[l][j].",
         " */",
         "#ifndef VLXMI_KERNELS_H",
         "#define VLXMI_KERNELS_H",
         "#include <stdint.h>",
         "#ifdef __cplusplus",
         'extern "C" {',
         "#endif",
         ""]
    for ab, cd in j_classes():
        cls = j_name(ab, cd)
        sig = emit._strip_restrict(emit.j_signature(cls))
        o.append("/* replaces computeCoulombFock%s, gen_coulomb.py:74-93 */" % cls.upper())
        o.append("int vlxmi_launch_computeCoulombFock%s(%s,\n    uint32_t ket_split, void* stream);" % (cls.upper(), ",\n    ".join(sig)))
        o.append("")
    for cls in k_classes():
        sig = emit._strip_restrict(emit.k_signature(cls))
        o.append("/* replaces computeExchangeFock%s, gen_exchange.py:75-115 */" % cls.upper())
        o.append("int vlxmi_launch_computeExchangeFock%s(%s,\n    void* stream);" % (cls.upper(), ",\n    ".join(sig)))
        o.append("")
    for ab, cd in j_classes():
        cls = j_name(ab, cd)
        sig = emit._strip_restrict(emit.j_signature(cls) + emit.MIXED_ARGS)
        o.append("/* mixed-precision variant of computeCoulombFock%s (no counterpart in the reference): quartets" % cls.upper())
        o.append(" * whose bound Qmax_IJ*max|Q_kl D_kl| is below fp32_threshold for every lane of a wavefront run the")
        o.append(" * VRR/HRR in FP32; Boys, prefactors and accumulation stay FP64.  quartet_counts (may be NULL):")
        o.append(" * [0] += primitive quartets evaluated in FP64, [1] += in FP32 (device, caller-zeroed) */")
        o.append("int vlxmi_launch_computeCoulombFock%s_mixed(%s,\n    uint32_t ket_split, void* stream);" % (cls.upper(), ",\n    ".join(sig)))
        o.append("")
    for cls in k_classes():
        sig = emit._strip_restrict(emit.k_signature(cls) + emit.MIXED_ARGS)
        o.append("/* mixed-precision variant of computeExchangeFock%s (no counterpart in the reference); bound" % cls.upper())
        o.append(" * QKmax_IJ*QKmax_KL*maxD_bd, quartet_counts as for the Coulomb variants */")
        o.append("int vlxmi_launch_computeExchangeFock%s_mixed(%s,\n    void* stream);" % (cls.upper(), ",\n    ".join(sig)))
        o.append("")
    o.append("/* Batched (persistent) variants: one launch covers many molecules whose lists are")
    o.append(" * concatenated.  J: work = n items of 4 uint32 (bra block begin, end <= begin+64, ket block")
    o.append(" * begin, end); each bra tile in at most one item.  K: blk_list = (i,k) blocks to process,")
    o.append(" * blk_mol = molecule of every block, mol_maxD/mol_D_off/mol_naos/mol_ao_off per molecule.")
    o.append(" * quartet_counts (may be NULL): [0] += primitive quartets evaluated in FP64, [1] += in FP32.")
    o.append(" * n_workgroups = persistent grid size (0: one per item).  No counterpart in the reference. */")
    for mixed in (False, True):
        for ab, cd in j_classes():
            cls = j_name(ab, cd)
            sig = emit._strip_restrict(emit.j_signature(cls) + ["const uint32_t* work", "const uint32_t nwork"]
                                       + emit.tail_args(mixed, True))
            o.append("int vlxmi_launch_computeCoulombFock%s_batched%s(%s,\n    uint32_t n_workgroups, void* stream);"
                     % (cls.upper(), "_mixed" if mixed else "", ",\n    ".join(sig)))
            o.append("")
        for cls in k_classes():
            sig = emit._strip_restrict(emit.k_signature(cls)) + [
                "const uint32_t* blk_list", "const uint32_t nlist", "const uint32_t* blk_mol",
                "const double* mol_maxD", "const unsigned long long* mol_D_off", "const uint32_t* mol_naos",
                "const uint32_t* mol_ao_off"] + emit._strip_restrict(emit.tail_args(mixed, True))
            o.append("int vlxmi_launch_computeExchangeFock%s_batched%s(%s,\n    uint32_t n_workgroups, void* stream);"
                     % (cls.upper(), "_mixed" if mixed else "", ",\n    ".join(sig)))
            o.append("")
    o.append("/* Launchers named after the reference's split kernels (VeloxChem launches these classes as")
    o.append(" * series computeCoulombFockPDDD0..6 etc., summary_fock_split.txt:1-20, all into one output):")
    o.append(" * split 0 runs the whole class, splits 1..n-1 launch nothing, so the series' sum is the class. */")
    for kind, series in SPLIT_SERIES.items():
        for cls, n in series.items():
            if kind == "J":
                sig = emit._strip_restrict(emit.j_signature(cls)) + ["uint32_t ket_split", "void* stream"]
                base = "computeCoulombFock" + cls.upper()
            else:
                sig = emit._strip_restrict(emit.k_signature(cls)) + ["void* stream"]
                base = "computeExchangeFock" + cls.upper()
            for k in range(n):
                o.append("int vlxmi_launch_%s%d(%s);" % (base, k, ", ".join(sig)))
    o.append("")
    o.append("/* Energy gradients at fixed density (no counterpart kernel in the reference; the derivative")
    o.append(" * operator is its Obara-Saika eq. 5, eri.py:85-175, write_eri_code(gradient=...) at")
    o.append(" * eriwriter.py:72-86).  Same lists and screening as the Fock launchers.")
    o.append(" * Coulomb: grad_ab[bb*6 + 0..5] += the A-centre (x,y,z) and B-centre parts of d/dX of")
    o.append(" *   E_J = 1/2 tr(D J) for every bra block bb; ab_mat_D_local = the bra entries' D (0 below the")
    o.append(" *   diagonal of a diagonal block, like the ket's); ket_split as for the Fock launchers (the")
    o.append(" *   chunks are added in a fixed order).")
    o.append(" * Exchange: grad_ac[blk*6 + 0..5] += the A-centre and C-centre parts of d/dX of")
    o.append(" *   E_K = 1/2 tr(D K) for every (i,k) block (factor 2 for blocks standing for (I,K) and (K,I)).")
    o.append(" * D must be symmetric.  Summing the rows into atoms: vlxmi/grad.py. */")
    for ab, cd in j_classes():
        cls = j_name(ab, cd)
        sig = emit._strip_restrict(emit_grad.jg_signature(cls))
        o.append("int vlxmi_launch_computeCoulombGradient%s(%s,\n    uint32_t ket_split, void* stream);" % (cls.upper(), ",\n    ".join(sig)))
        o.append("")
    for cls in k_classes():
        sig = emit._strip_restrict(emit_grad.kg_signature(cls))
        o.append("int vlxmi_launch_computeExchangeGradient%s(%s,\n    void* stream);" % (cls.upper(), ",\n    ".join(sig)))
        o.append("")
    for pc in PAIR_CLASSES:
        o.append("/* Schwarz factors sqrt|(ab|ab)| per component pair of primitive-pair blocks (the")
        o.append(" * reference's mat_Q / Q_K inputs, computed by VeloxChem's host there). */")
        o.append("int vlxmi_schwarz_%s(const double* a_info, uint32_t a_count, const double* b_info, uint32_t b_count,\n"
                 "    const uint32_t* Iblk, const uint32_t* Jblk, const double* S, uint32_t nblk, double* Qout,\n"
                 "    const double* boys_func_table, void* stream);" % pc)
        o.append("")
    o += ["#ifdef __cplusplus", "}", "#endif", "#endif", ""]
    return "\n".join(o)


def _compile(src_path, obj_path, extra=()):
    cmd = [HIPCC] + FLAGS + list(extra) + ["-c", src_path, "-o", obj_path + ".tmp"]
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src_path, r.stderr[-4000:]))
    os.replace(obj_path + ".tmp", obj_path)
    return time.time() - t


def kernel_resources(obj_path):
    """{kernel name: (registers per lane, LDS bytes)} of the gfx950 code object
    inside a host object compiled by this module."""
    with tempfile.TemporaryDirectory() as td:
        fb, dv = os.path.join(td, "fb.bin"), os.path.join(td, "dev.o")
        subprocess.run([LLVM + "/llvm-objcopy", "--dump-section", ".hip_fatbin=" + fb, obj_path], check=True)
        subprocess.run([LLVM + "/clang-offload-bundler", "--unbundle", "--type=o", "--input=" + fb,
                        "--output=" + dv, "--targets=hipv4-amdgcn-amd-amdhsa--" + ARCH], check=True)
        txt = subprocess.run([LLVM + "/llvm-readelf", "--notes", dv], capture_output=True, text=True,
                             check=True).stdout
    out = {}
    for blk in txt.split("- .agpr_count")[1:]:
        name = re.search(r"\.name:\s+(\S+)", blk).group(1)
        regs = int(re.search(r"\.vgpr_count:\s+(\d+)", blk).group(1))
        lds = int(re.search(r"\.group_segment_fixed_size:\s+(\d+)", blk).group(1))
        out[name] = (regs, lds)
    return out


def kernel_scratch(obj_path):
    """{kernel name: scratch bytes per lane (private segment)} of the gfx950
    code object inside a host object compiled by this module."""
    with tempfile.TemporaryDirectory() as td:
        fb, dv = os.path.join(td, "fb.bin"), os.path.join(td, "dev.o")
        subprocess.run([LLVM + "/llvm-objcopy", "--dump-section", ".hip_fatbin=" + fb, obj_path], check=True)
        subprocess.run([LLVM + "/clang-offload-bundler", "--unbundle", "--type=o", "--input=" + fb,
                        "--output=" + dv, "--targets=hipv4-amdgcn-amd-amdhsa--" + ARCH], check=True)
        txt = subprocess.run([LLVM + "/llvm-readelf", "--notes", dv], capture_output=True, text=True,
                             check=True).stdout
    out = {}
    for blk in txt.split("- .agpr_count")[1:]:
        name = re.search(r"\.name:\s+(\S+)", blk).group(1)
        out[name] = int(re.search(r"\.private_segment_fixed_size:\s+(\d+)", blk).group(1))
    return out


def tile_for(regs, lds, kt, rec_bytes, min_kt=4):
    """Largest power-of-two ket tile <= kt whose LDS lets the kernel reach the
    waves per CU its register count allows."""
    granule = -(-max(regs, 1) // 8) * 8
    waves_cu = 4 * min(8, 512 // granule)
    cap = 160 * 1024 // waves_cu
    fixed = lds - kt * rec_bytes
    t = kt
    while t > min_kt and fixed + t * rec_bytes > cap:
        t //= 2
    return t


def retune(objs_by_src, tiles):
    """New tile table from the compiled kernels' resources."""
    new = dict(tiles)
    for fn, obj in objs_by_src.items():
        m = re.match(r"(j|jm|jb|jbm|k|km|kb|kbm)_([spd]{4})\.hip$", fn)
        if not m:
            continue
        kind = "J" if m.group(1).startswith("j") else "K"
        cls = m.group(2)
        res = kernel_resources(obj)
        (regs, lds), = [v for k, v in res.items() if "Fock" in k]
        src = open(os.path.join(GEN, fn)).read()
        kt = int(re.search(r"constexpr int KT = (\d+)", src).group(1))
        t = tile_for(regs, lds, kt, emit.ket_record_bytes(kind, cls))
        if t != kt:
            new[fn] = t
    return new


def build(jobs=None, force=False, only=None, verbose=True, retune_tiles=False):
    lib = _build(jobs, force, only, verbose)
    if retune_tiles:
        tiles = load_tiles()
        new = retune(_last_objs, tiles)
        if new != tiles:
            with open(TILE_TABLE, "w") as f:
                json.dump(dict(sorted(new.items())), f, indent=1)
                f.write("\n")
            if verbose:
                print("ket tiles retuned for %d kernels; rebuilding" % len(new), flush=True)
            lib = _build(jobs, False, only, verbose)
    return lib


_last_objs = {}


def _build(jobs=None, force=False, only=None, verbose=True):
    os.makedirs(GEN, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    with open(HEADER, "w") as f:
        f.write(header_text())
    with open(DEVICE_HEADER, "w") as f:
        f.write(device_header_text())
    todo = []
    objs = []
    hand = ["vlxmi_capi.cpp", "vlxmi_noalloc.cpp", "vlxmi_lists.cpp", "vlxmi_pairs.cpp"]
    flag_key = " ".join(FLAGS)
    dev_hdr = open(os.path.join(CSRC, "vlxmi_device.hpp")).read() + open(os.path.join(CSRC, "vlxmi_status.h")).read()
    for fn in hand:
        p = os.path.join(CSRC, fn)
        text = open(p).read()
        h = hashlib.sha1((text + dev_hdr + flag_key).encode()).hexdigest()[:16]
        o = os.path.join(OBJ, "%s.%s.o" % (fn, h))
        objs.append(o)
        if force or not os.path.exists(o):
            todo.append((p, o, ("-x", "hip")))
    for fn, text in srcs.items():
        p = os.path.join(GEN, fn)
        old = open(p).read() if os.path.exists(p) else None
        if old != text:
            with open(p, "w") as f:
                f.write(text)
        h = hashlib.sha1((text + dev_hdr + flag_key).encode()).hexdigest()[:16]
        o = os.path.join(OBJ, "%s.%s.o" % (fn, h))
        objs.append(o)
        _last_objs[fn] = o
        if only and not any(x in fn for x in only):
            continue
        if force or not os.path.exists(o):
            todo.append((p, o, ()))
    jobs = jobs or max(1, min(8, os.cpu_count() or 1))
    t0 = time.time()
    # compile the heaviest first (d-classes)
    todo.sort(key=lambda x: -os.path.getsize(x[0]))
    if todo:
        with cf.ThreadPoolExecutor(jobs) as ex:
            futs = {ex.submit(_compile, p, o, e): p for p, o, e in todo}
            for fu in cf.as_completed(futs):
                dt = fu.result()
                if verbose:
                    print("  built %-40s %6.1fs" % (os.path.basename(futs[fu]), dt), flush=True)
    missing = [o for o in objs if not os.path.exists(o)]
    if missing:
        raise RuntimeError("missing objects (use only=None): %d" % len(missing))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    # drop stale objects
    keep = set(objs)
    for fn in os.listdir(OBJ):
        p = os.path.join(OBJ, fn)
        if p not in keep and fn.endswith(".o"):
            os.remove(p)
    if verbose:
        print("libvlxmi.so: %d objects (%d compiled) in %.1fs" % (len(objs), len(todo), time.time() - t0), flush=True)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--only", nargs="*")
    ap.add_argument("--retune", action="store_true",
                    help="re-derive the ket-tile table from the compiled kernels' registers and LDS")
    a = ap.parse_args(argv)
    build(jobs=a.jobs, force=a.force, only=a.only, retune_tiles=a.retune)


if __name__ == "__main__":
    main()
