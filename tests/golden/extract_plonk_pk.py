"""Write the golden vectors held in the reference's cached PLONK proving key.

The reference ships one serialized file, a PLONK proving key for a domain of
n = 16 written by its setup cache (runtime/cache.py:70-79):
``.pyZKP_cache/setup/plonk/ir_3d1d5f0ed874c6e1_mod2188...617_v1.pkl``.

That file is never unpickled here.  ``pickletools.genops`` only parses the
opcode stream, and the tiny stack machine below rebuilds plain data from it
(ints, strings, lists, tuples, dicts).  Class references become inert
``{"__class__": "module.Name", ...state}`` dicts: nothing is imported, no
constructor or ``__setstate__`` runs, no code from the file executes.

Output: ``plonk_pk_n16.json`` beside this script, holding
  * n, omega (canonical ints),
  * the first 16 SRS G1 points (affine, canonical x/y),
  * the 8 selector/permutation evaluation vectors and their 8 inverse-NTT
    coefficient vectors (length 16, canonical),
  * the 8 KZG commitments (affine), i.e. 8 G1 MSMs of 16 SRS points,
  * the first two G2 SRS points (affine, Fq2 as [c0, c1]).

Run from the repository root:  python tests/golden/extract_plonk_pk.py
(it reads the reference checkout at /root/reference, which only this
generator needs; the JSON it writes is what the tests load).
"""
from __future__ import annotations

import json
import os
import pickletools
import sys

PKL = ("/root/reference/.pyZKP_cache/setup/plonk/"
       "ir_3d1d5f0ed874c6e1_mod21888242871839275222246405745257275088548364400416034343698204186575808495617_v1.pkl")
Q = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47


class _Mark:
    pass


_MARK = _Mark()


def _inert_load(data: bytes):
    """Rebuild plain data from a pickle opcode stream without executing it."""
    stack: list = []
    memo: dict = {}
    for op, arg, _pos in pickletools.genops(data):
        name = op.name
        if name in ("PROTO", "FRAME"):
            continue
        if name == "STOP":
            break
        if name == "MARK":
            stack.append(_MARK)
        elif name == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif name in ("BINGET", "LONG_BINGET", "GET"):
            stack.append(memo[arg])
        elif name in ("BINPUT", "LONG_BINPUT", "PUT"):
            memo[arg] = stack[-1]
        elif name in ("BININT", "BININT1", "BININT2", "LONG1", "LONG4", "INT", "LONG"):
            stack.append(int(arg))
        elif name in ("SHORT_BINUNICODE", "BINUNICODE", "BINUNICODE8", "UNICODE"):
            stack.append(str(arg))
        elif name in ("NEWTRUE", "NEWFALSE"):
            stack.append(name == "NEWTRUE")
        elif name == "NONE":
            stack.append(None)
        elif name == "EMPTY_DICT":
            stack.append({})
        elif name == "EMPTY_LIST":
            stack.append([])
        elif name == "EMPTY_TUPLE":
            stack.append(())
        elif name in ("TUPLE1", "TUPLE2", "TUPLE3"):
            k = int(name[-1])
            items = tuple(stack[-k:])
            del stack[-k:]
            stack.append(items)
        elif name == "TUPLE":
            i = max(j for j, v in enumerate(stack) if v is _MARK)
            items = tuple(stack[i + 1:])
            del stack[i:]
            stack.append(items)
        elif name == "APPEND":
            v = stack.pop()
            stack[-1].append(v)
        elif name == "APPENDS":
            i = max(j for j, v in enumerate(stack) if v is _MARK)
            items = stack[i + 1:]
            del stack[i:]
            stack[-1].extend(items)
        elif name == "SETITEM":
            v = stack.pop()
            k = stack.pop()
            stack[-1][k] = v
        elif name == "SETITEMS":
            i = max(j for j, v in enumerate(stack) if v is _MARK)
            items = stack[i + 1:]
            del stack[i:]
            d = stack[-1]
            for a in range(0, len(items), 2):
                d[items[a]] = items[a + 1]
        elif name == "STACK_GLOBAL":
            nm = stack.pop()
            mod = stack.pop()
            stack.append(("__global__", f"{mod}.{nm}"))
        elif name == "NEWOBJ":
            _args = stack.pop()
            cls = stack.pop()
            stack.append({"__class__": cls[1]})
        elif name == "BUILD":
            state = stack.pop()
            obj = stack[-1]
            if isinstance(state, dict):
                obj.update(state)
            else:
                obj["__state__"] = state
        else:
            raise ValueError(f"opcode {name} not handled by the inert loader")
    assert len(stack) == 1
    return stack[0]


def _fq(v) -> int:
    if isinstance(v, dict):  # py_ecc FQ: state {'n': int}
        return int(v["n"])
    return int(v)


def _g1_affine(p):
    x, y, z = (_fq(c) for c in p)
    if z == 0:
        return None
    zi = pow(z, -1, Q)  # py_ecc optimized_bn128 is homogeneous projective
    return [x * zi % Q, y * zi % Q]


def _fq2(v):
    if isinstance(v, dict):
        c = v.get("coeffs")
        return [int(c[0]), int(c[1])]
    raise ValueError("unexpected FQ2 encoding")


def _fq2_mul(a, b):
    return [(a[0] * b[0] - a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q]


def _fq2_inv(a):
    t = pow((a[0] * a[0] + a[1] * a[1]) % Q, -1, Q)
    return [a[0] * t % Q, (-a[1]) * t % Q]


def _g2_affine(p):
    x, y, z = (_fq2(c) for c in p)
    if z == [0, 0]:
        return None
    zi = _fq2_inv(z)
    return [_fq2_mul(x, zi), _fq2_mul(y, zi)]


def main() -> None:
    with open(PKL, "rb") as f:
        root = _inert_load(f.read())
    meta = root["meta"]
    pk = root["data"]
    dom = pk["circuit"]["domain"]
    n = int(dom["n"])
    omega = int(dom["omega"])
    srs = pk["srs"]
    names = ["ql", "qr", "qm", "qo", "qc", "sigma1", "sigma2", "sigma3"]
    circ = pk["circuit"]
    vk = pk["vk"]
    evals = {nm: [int(x) for x in circ[f"{nm}_eval"]] for nm in names}
    coeffs = {nm: [int(x) for x in pk[f"coeff_{nm}"]] for nm in names}
    commits = {nm: _g1_affine(vk[f"cm_{nm}"]) for nm in names}
    g1 = srs["g1_powers"]
    g2 = srs["g2_powers"]
    out = {
        "source": "reference .pyZKP_cache PLONK proving key (runtime/cache.py:70-79), decoded inertly",
        "meta": {k: (str(v) if not isinstance(v, int) else v) for k, v in meta.items()},
        "n": n,
        "omega": omega,
        "roots": [int(x) for x in dom["roots"]],
        "srs_g1_len": len(g1),
        "srs_g2_len": len(g2),
        "srs_g1": [_g1_affine(p) for p in g1[:n]],
        "srs_g2": [_g2_affine(p) for p in g2[:2]],
        "evals": evals,
        "coeffs": coeffs,
        "commits": commits,
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plonk_pk_n16.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst, file=sys.stderr)


if __name__ == "__main__":
    main()
