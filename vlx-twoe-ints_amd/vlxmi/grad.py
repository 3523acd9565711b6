"""Coulomb and exchange energy gradients on one GPU.

``GradientEngine(eng)`` runs the ``computeCoulombGradient{AB}{CD}`` and
``computeExchangeGradient{A}{B}{C}{D}`` kernels over a ``FockEngine``'s
resident lists and sums their per-block rows into per-atom gradients:

  gJ[X] = d/dX of E_J = 1/2 sum_ab D_ab J_ab[D]
  gK[X] = d/dX of E_K = 1/2 sum_ab D_ab K_ab[D]

at fixed density D (the two-electron part of an SCF gradient; the
energy-weighted-density and one-electron terms belong to the caller).  A
closed-shell Hartree-Fock energy with the total density D has the
two-electron part E_J - E_K / 2, so its gradient is gJ - gK / 2.  The
integral-derivative operator is the reference's (Obara-Saika eq. 5,
eri.py:85-175); the screening is the Fock build's, so the quartets that
contribute are those the J and K kernels evaluate.

Each Coulomb bra block yields six numbers (the A-centre and B-centre parts),
each exchange (I,K) block six (A-centre and C-centre parts); the per-atom sums
are one deterministic segmented sum (plans sorted once), so the gradient is
reproducible bit for bit.
"""

from __future__ import annotations

import numpy as np

from . import pairs as P
from .assemble import Assembler
from .basis import NCART
from .runtime import check, lib


def prim_atoms(fl):
    """Atom of every primitive, per angular momentum (None where absent)."""
    shell_atom = np.array([s.atom for s in fl.mol.shells], dtype=np.int64)
    return [shell_atom[p.shell] if p is not None else None for p in fl.prims]


def j_block_atoms(fl, ab, patoms=None):
    """(atom of a, atom of b) of every primitive-pair block of the J list ``ab``."""
    patoms = patoms or prim_atoms(fl)
    jl = fl.jlists[ab]
    nab = NCART[jl.la] * NCART[jl.lb]
    first = jl.first[::nab].astype(np.int64) // NCART[jl.la]
    second = jl.second[::nab].astype(np.int64) // NCART[jl.lb]
    return patoms[jl.la][first], patoms[jl.lb][second]


def k_block_atoms(fl, ac, patoms=None):
    """(atom of a, atom of c) of every (I,K) block of the (i,k) list ``ac``."""
    patoms = patoms or prim_atoms(fl)
    ik = fl.iklists[ac]
    return patoms[ik.la][ik.I.astype(np.int64)], patoms[ik.lc][ik.K.astype(np.int64)]


def row_map(atoms_a, atoms_b, offset):
    """(value index, output index) pairs sending rows [nblk, 6] (first centre
    x, y, z, second centre x, y, z) at ``offset`` to per-atom outputs atom*3+g."""
    nblk = atoms_a.size
    vi = offset + np.arange(6 * nblk, dtype=np.int64)
    at = np.stack([atoms_a] * 3 + [atoms_b] * 3, axis=1)
    g = np.tile(np.array([0, 1, 2, 0, 1, 2], dtype=np.int64), (nblk, 1))
    return vi, (3 * at + g).reshape(-1)


def sum_rows(natoms, parts):
    """Host version of the per-atom sum: parts = [(rows [nblk, 6], atoms_a,
    atoms_b)] -> (natoms, 3)."""
    g = np.zeros((natoms, 3))
    for rows, aa, bb in parts:
        np.add.at(g, aa, rows[:, :3])
        np.add.at(g, bb, rows[:, 3:])
    return g


class GradientEngine:
    def __init__(self, eng):
        self.eng = eng
        self.dev = eng.dev
        self.torch = eng.torch
        fl = eng.fl
        self.natoms = fl.mol.natoms
        pa = prim_atoms(fl)
        vi, oi, off = [], [], 0
        self.j_atoms = {}
        for pc in eng.j_order:
            aa, bb = j_block_atoms(fl, pc, pa)
            self.j_atoms[pc] = (aa, bb)
            v, o = row_map(aa, bb, off)
            vi.append(v)
            oi.append(o)
            off += 6 * aa.size
        self.jrows = off
        self.jasm = Assembler(self.dev, np.concatenate(vi) if vi else np.zeros(0), np.concatenate(oi) if oi else np.zeros(0),
                              3 * self.natoms)
        vi, oi, off = [], [], 0
        self.k_atoms = {}
        for pc in eng.k_order:
            aa, cc = k_block_atoms(fl, pc, pa)
            self.k_atoms[pc] = (aa, cc)
            v, o = row_map(aa, cc, off)
            vi.append(v)
            oi.append(o)
            off += 6 * aa.size
        self.krows = off
        self.kasm = Assembler(self.dev, np.concatenate(vi) if vi else np.zeros(0), np.concatenate(oi) if oi else np.zeros(0),
                              3 * self.natoms)

    # ------------------------------------------------------------------ launches
    def run_J(self, Dc, kets=None, classes=None, bra_blocks=None):
        """Per bra class the gradient rows [blocks, 6] (A-centre, B-centre
        parts), summed over ket classes; ``bra_blocks`` ({ab: block indices})
        runs each bra class on that slice of its list (rows in slice order)."""
        e, t, L = self.eng, self.torch, lib()
        kets = kets or e.prepare_kets(Dc)
        out, slices = {}, {}
        for ab, cd in (classes or P.j_classes()):
            if ab not in e.jl or cd not in e.jl:
                continue
            if ab not in slices:
                if bra_blocks is not None:
                    s = e.bra_slice(ab, bra_blocks[ab])
                    nab = s["nab"]
                    b = t.as_tensor(np.asarray(bra_blocks[ab], dtype=np.int64), device=self.dev.device)
                    ent = (b[:, None] * nab + t.arange(nab, device=self.dev.device)[None, :]).reshape(-1)
                    s["mat_D"] = kets[ab]["mat_D_bra"][ent].contiguous()
                else:
                    s = dict(e.jl[ab], mat_D=kets[ab]["mat_D_bra"])
                slices[ab] = s
            bra, ket = slices[ab], kets[cd]
            if ab not in out:
                out[ab] = t.zeros((bra["nblk"], 6), dtype=t.float64, device=Dc.device)
            cls = ab + cd
            args = [out[ab].data_ptr()]
            for x in "spd":
                if x in cls:
                    l = P.LVAL[x]
                    args += [e.prim_info[l].data_ptr(), e.prim_count[l]]
            args += [bra["mat_D"].data_ptr(), ket["mat_D"].data_ptr(), bra["mat_Q"].data_ptr(), ket["mat_Q"].data_ptr(),
                     bra["first"].data_ptr(), bra["second"].data_ptr(), bra["pair_data"].data_ptr(), bra["count"],
                     ket["first"].data_ptr(), ket["second"].data_ptr(), ket["pair_data"].data_ptr(), ket["count"],
                     self.dev.boys_table.data_ptr(), self.dev.boys_ft.data_ptr(), e.thr]
            split = e._ket_split(bra["nblk"], e.jl[cd]["nblk"])
            name = "computeCoulombGradient" + cls.upper()
            ev = e._ev_start()
            check(getattr(L, "vlxmi_launch_" + name)(*args, split, self.dev.stream_ptr()), name)
            e._ev_end(ev, name)
        return out

    def run_K(self, Dc, classes=None, maxD=None, ik_blocks=None):
        """Per (a,c) class the gradient rows [blocks, 6] (A-centre, C-centre
        parts), summed over (b,d); ``ik_blocks`` as in FockEngine.run_K."""
        e, t, L = self.eng, self.torch, lib()
        maxD = maxD or e.max_D(Dc)
        out, slices = {}, {}
        for cls in (classes or P.k_classes()):
            a, b, c, d = cls
            ac, ab, cd, bd = a + c, a + b, c + d, b + d
            if ac not in e.ik or ab not in e.kc or cd not in e.kc:
                continue
            if ik_blocks is not None and ac not in slices:
                slices[ac] = e.ik_slice(ac, ik_blocks[ac])
            ik = slices.get(ac, e.ik[ac])
            if ac not in out:
                out[ac] = t.zeros((ik["nblk"], 6), dtype=t.float64, device=Dc.device)
            args = [out[ac].data_ptr(), ik["i"].data_ptr(), ik["k"].data_ptr(), ik["count"]]
            for x in "spd":
                if x in cls:
                    l = P.LVAL[x]
                    args += [e.prim_info[l].data_ptr(), e.prim_aoinds[l].data_ptr(), e.prim_count[l]]
            args += [maxD[bd], Dc.data_ptr(), e.naos]
            A, Cc = e.kc[ab], e.kc[cd]
            for nm in ("Q", "D_inds", "displs", "counts", "pair_data"):
                args.append(A[nm].data_ptr())
                if cd != ab:
                    args.append(Cc[nm].data_ptr())
            args += [self.dev.boys_table.data_ptr(), self.dev.boys_ft.data_ptr(), e.omega, e.thr]
            name = "computeExchangeGradient" + cls.upper()
            ev = e._ev_start()
            check(getattr(L, "vlxmi_launch_" + name)(*args, self.dev.stream_ptr()), name)
            e._ev_end(ev, name)
        return out

    # ------------------------------------------------------------------ assembly
    def _rows(self, m, order, nblk):
        t = self.torch
        parts = [m[pc].reshape(-1) if pc in m else t.zeros(6 * nblk[pc], dtype=t.float64, device=self.dev.device)
                 for pc in order]
        return t.cat(parts) if parts else t.zeros(0, dtype=t.float64, device=self.dev.device)

    def assemble(self, rows, kind):
        """Per-atom (natoms, 3) gradient from the per-class rows of run_J
        ("J") or run_K ("K") over the whole lists."""
        t = self.torch
        e = self.eng
        if kind == "J":
            vals = self._rows(rows, e.j_order, {pc: e.jl[pc]["nblk"] for pc in e.j_order})
            asm = self.jasm
        else:
            vals = self._rows(rows, e.k_order, {pc: e.ik[pc]["nblk"] for pc in e.k_order})
            asm = self.kasm
        g = t.zeros(3 * self.natoms, dtype=t.float64, device=self.dev.device)
        if vals.numel():
            asm.run(vals, g)
        return g.view(self.natoms, 3)

    def gradient(self, D_sph, want_k=True):
        """(gJ, gK) per atom (device tensors, (natoms, 3)) for density D."""
        Dc = self.eng.cart_density(D_sph)
        gJ = self.assemble(self.run_J(Dc), "J")
        gK = self.assemble(self.run_K(Dc), "K") if want_k else None
        return gJ, gK
