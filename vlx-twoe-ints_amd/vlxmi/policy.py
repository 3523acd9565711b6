"""Which classes run their ``_mixed`` kernel in mixed precision.

gfx950 issues FP32 VALU instructions at the FP64 rate when packed FP32 is not
used (the library is compiled without it), so running a class's VRR/HRR in FP32
only pays where halving the recursion's registers removes spills; in the light
classes the conversions make the ``_mixed`` kernel slower than the FP64 one.
``gen/mixed_classes.json`` lists the classes whose ``_mixed`` kernel measured
at least 3% faster than their FP64 kernel on lattice64; with the "measured"
policy the engines launch the ``_mixed`` kernel for those classes and the FP64
kernel for every other class, with policy "all" every class runs ``_mixed``.
Either way no quartet above the FP32 bound is evaluated in FP32.
"""

from __future__ import annotations

import json
import os

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gen", "mixed_classes.json")
POLICIES = ("measured", "all")
_cache = None


def mixed_classes():
    """{"J": set of 'abcd', "K": set of 'abcd'} from the measured table."""
    global _cache
    if _cache is None:
        with open(TABLE) as f:
            tab = json.load(f)
        _cache = {"J": set(tab["J"]), "K": set(tab["K"])}
    return _cache


def use_mixed(kind: str, cls: str, precision: str, policy: str = "measured") -> bool:
    """True when class ``cls`` (kind "J" or "K") launches its _mixed kernel."""
    if precision != "mixed":
        return False
    if policy not in POLICIES:
        raise ValueError("mixed policy must be one of %s" % (POLICIES,))
    return policy == "all" or cls in mixed_classes()[kind]
