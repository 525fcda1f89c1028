"""Basis-set data tables in Basis Set Exchange JSON layout.

The reference fetches its basis data from the ``basis_set_exchange`` package
(jax_qc/basis/bse_fetch.py:18-40), which is not installed here.  The tables
below are typed in from the published definitions, in the same BSE dict
layout (``elements -> Z -> electron_shells -> {angular_momentum, exponents,
coefficients}``, numbers as strings), with combined sp shells left combined:
the parser splits them, as ``uncontract_spdf=True`` does for the reference.

* STO-3G: Hehre, Stewart & Pople, J. Chem. Phys. 51, 2657 (1969).
* 6-31G: Hehre, Ditchfield & Pople, J. Chem. Phys. 56, 2257 (1972).
* 6-31G** polarisation (d on C-F: 0.8; p on H: 1.1): Hariharan & Pople,
  Theor. Chim. Acta 28, 213 (1973).

The example energies the reference hard-codes (examples/07_rhf_scf.py:31-36,
examples/08_rhf_tier2.py:33-41) are the acceptance test for these numbers.
"""

from __future__ import annotations

import copy
from typing import Dict

_S1 = ["0.1543289673E+00", "0.5353281423E+00", "0.4446345422E+00"]
_S2 = ["-0.9996722919E-01", "0.3995128261E+00", "0.7001154689E+00"]
_P2 = ["0.1559162750E+00", "0.6076837186E+00", "0.3919573931E+00"]


def _sto3g_first_row(e1s, e2sp):
    return {
        "electron_shells": [
            {"angular_momentum": [0], "exponents": e1s, "coefficients": [_S1]},
            {"angular_momentum": [0, 1], "exponents": e2sp, "coefficients": [_S2, _P2]},
        ]
    }


STO_3G: Dict = {
    "name": "sto-3g",
    "elements": {
        "1": {
            "electron_shells": [
                {
                    "angular_momentum": [0],
                    "exponents": ["0.3425250914E+01", "0.6239137298E+00", "0.1688554040E+00"],
                    "coefficients": [_S1],
                }
            ]
        },
        "2": {
            "electron_shells": [
                {
                    "angular_momentum": [0],
                    "exponents": ["0.6362421394E+01", "0.1158922999E+01", "0.3136497915E+00"],
                    "coefficients": [_S1],
                }
            ]
        },
        "6": _sto3g_first_row(
            ["0.7161683735E+02", "0.1304509632E+02", "0.3530512160E+01"],
            ["0.2941249355E+01", "0.6834830964E+00", "0.2222899159E+00"],
        ),
        "7": _sto3g_first_row(
            ["0.9910616896E+02", "0.1805231239E+02", "0.4885660238E+01"],
            ["0.3780455879E+01", "0.8784966449E+00", "0.2857143744E+00"],
        ),
        "8": _sto3g_first_row(
            ["0.1307093214E+03", "0.2380886605E+02", "0.6443608313E+01"],
            ["0.5033151319E+01", "0.1169596125E+01", "0.3803889600E+00"],
        ),
        "9": _sto3g_first_row(
            ["0.1666791340E+03", "0.3036081233E+02", "0.8216820672E+01"],
            ["0.6464803249E+01", "0.1502281245E+01", "0.4885884864E+00"],
        ),
    },
}


def _631g_first_row(core_exp, core_coef, sp_exp, s_coef, p_coef, outer):
    return {
        "electron_shells": [
            {"angular_momentum": [0], "exponents": core_exp, "coefficients": [core_coef]},
            {"angular_momentum": [0, 1], "exponents": sp_exp, "coefficients": [s_coef, p_coef]},
            {
                "angular_momentum": [0, 1],
                "exponents": [outer],
                "coefficients": [["0.1000000000E+01"], ["0.1000000000E+01"]],
            },
        ]
    }


_631G_H = {
    "electron_shells": [
        {
            "angular_momentum": [0],
            "exponents": ["0.1873113696E+02", "0.2825394365E+01", "0.6401216923E+00"],
            "coefficients": [["0.3349460434E-01", "0.2347269535E+00", "0.8137573261E+00"]],
        },
        {
            "angular_momentum": [0],
            "exponents": ["0.1612777588E+00"],
            "coefficients": [["0.1000000000E+01"]],
        },
    ]
}

_631G_O = _631g_first_row(
    ["0.5484671660E+04", "0.8252349460E+03", "0.1880469580E+03", "0.5296450000E+02",
     "0.1689757040E+02", "0.5799635340E+01"],
    ["0.1831074430E-02", "0.1395017220E-01", "0.6844507810E-01", "0.2327143360E+00",
     "0.4701928980E+00", "0.3585208530E+00"],
    ["0.1553961625E+02", "0.3599933586E+01", "0.1013761750E+01"],
    ["-0.1107775495E+00", "-0.1480262627E+00", "0.1130767015E+01"],
    ["0.7087426823E-01", "0.3397528391E+00", "0.7271585773E+00"],
    "0.2700058226E+00",
)

_631G_F = _631g_first_row(
    ["0.7001713090E+04", "0.1051366090E+04", "0.2392856900E+03", "0.6739744530E+02",
     "0.2151995730E+02", "0.7403101300E+01"],
    ["0.1819616900E-02", "0.1391607960E-01", "0.6840532450E-01", "0.2331857600E+00",
     "0.4712674390E+00", "0.3566185460E+00"],
    ["0.2084795280E+02", "0.4808308340E+01", "0.1344069860E+01"],
    ["-0.1085069750E+00", "-0.1464516580E+00", "0.1128688580E+01"],
    ["0.7162872430E-01", "0.3459121030E+00", "0.7224699570E+00"],
    "0.3581513930E+00",
)

SIX_31G: Dict = {"name": "6-31g", "elements": {"1": _631G_H, "8": _631G_O, "9": _631G_F}}


def _with_polarisation(elem: Dict, l: int, exponent: str) -> Dict:
    out = copy.deepcopy(elem)
    out["electron_shells"].append(
        {"angular_momentum": [l], "exponents": [exponent], "coefficients": [["0.1000000000E+01"]]}
    )
    return out


SIX_31GSS: Dict = {
    "name": "6-31g**",
    "elements": {
        "1": _with_polarisation(_631G_H, 1, "0.1100000000E+01"),
        "8": _with_polarisation(_631G_O, 2, "0.8000000000E+00"),
        "9": _with_polarisation(_631G_F, 2, "0.8000000000E+00"),
    },
}

LIBRARY = {"sto-3g": STO_3G, "6-31g": SIX_31G, "6-31g**": SIX_31GSS}


def get_basis_dict(name: str) -> Dict:
    key = name.strip().lower()
    if key not in LIBRARY:
        raise KeyError(f"basis {name!r} not in the built-in library ({sorted(LIBRARY)})")
    return LIBRARY[key]
