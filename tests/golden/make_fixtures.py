"""Write tests/golden/reference_vectors.json from the reference's own files.

The reference is READ here, never imported or executed: every constant is
pulled out of the Python sources with `ast` and evaluated by the small
arithmetic evaluator below (numbers, + - * // % ** <<, names bound earlier in
the same file, and the reference's `_down` alignment helper), and the
evidence ledgers are loaded with `json`.  What lands in the fixture is data:
frozen benchmark fingerprints, literature tables, measured gate windows and
discovery records, each tagged with the file:line it came from.

Run from the repository root:  python tests/golden/make_fixtures.py
(needs the reference tree at /root/reference; the fixture it writes is
committed, so nothing else ever needs the reference).
"""

import ast
import json
import operator
import pathlib
import re
import sys

REF = pathlib.Path("/root/reference")
OUT = pathlib.Path(__file__).with_name("reference_vectors.json")

_BIN = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.FloorDiv: operator.floordiv, ast.Mod: operator.mod,
        ast.Pow: operator.pow, ast.LShift: operator.lshift}


def _eval(node, env):
    """Evaluate a constant arithmetic expression node against `env`."""
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, ast.Name):
        return env[node.id]
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -_eval(node.operand, env)
    if isinstance(node, ast.BinOp) and type(node.op) in _BIN:
        return _BIN[type(node.op)](_eval(node.left, env),
                                   _eval(node.right, env))
    if isinstance(node, (ast.Tuple, ast.List)):
        return [_eval(e, env) for e in node.elts]
    if isinstance(node, ast.Dict):
        return {_eval(k, env): _eval(v, env)
                for k, v in zip(node.keys, node.values)}
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) \
            and node.func.id in env and callable(env[node.func.id]):
        return env[node.func.id](*[_eval(a, env) for a in node.args])
    raise ValueError(f"not a constant expression: {ast.dump(node)[:80]}")


def module_constants(rel, env=None, names=None):
    """{name: (value, 'rel:line')} for top-level constant assignments."""
    env = dict(env or {})
    src = (REF / rel).read_text()
    out = {}
    for node in ast.parse(src).body:
        if not isinstance(node, ast.Assign) or len(node.targets) != 1:
            continue
        tgt = node.targets[0]
        if not isinstance(tgt, ast.Name):
            continue
        try:
            val = _eval(node.value, env)
        except (ValueError, KeyError, TypeError):
            continue
        env[tgt.id] = val
        if names is None or tgt.id in names:
            out[tgt.id] = (val, f"{rel}:{node.lineno}")
    return out, src


def _jsonable(v):
    if isinstance(v, int) and abs(v) >= 2**53:
        return str(v)                     # exact big ints travel as strings
    if isinstance(v, list):
        return [_jsonable(x) for x in v]
    if isinstance(v, dict):
        return {str(k): _jsonable(x) for k, x in v.items()}
    return v


def main():
    if not REF.is_dir():
        sys.exit("reference tree not mounted; the committed fixture stands")
    fx = {"_about": "Reference-held vectors, extracted by "
                    "tests/golden/make_fixtures.py (text + json only). "
                    "Integers >= 2^53 are decimal strings."}

    # ---- huntlib/primes.py: the strong-test bases and small primes
    c, _ = module_constants("huntlib/primes.py")
    fx["mr"] = {k: {"value": _jsonable(v), "at": at}
                for k, (v, at) in c.items()
                if k in ("MR_BASES", "MR_VALID_BELOW", "_SMALL")}

    # ---- Euler frozen benchmark fingerprints (n = 17, exact-p XOR)
    c, _ = module_constants("euler-prime-runs/score.py")
    shapes = []
    for pre, label in (("BENCH", "SCORE"), ("BENCH128", "SCORE128"),
                       ("BENCH_WIDE", "SCORE_WIDE")):
        fp = "FINGERPRINT" + pre[len("BENCH"):]
        lo, at = c[pre + "_LO"]
        shapes.append({"label": label, "n": 17, "lo": _jsonable(lo),
                       "span": _jsonable(c[pre + "_SPAN"][0]),
                       "count": c[fp + "_COUNT"][0],
                       "xor": _jsonable(c[fp + "_CHECKSUM"][0]),
                       "at": at})
    fx["euler_fingerprints"] = shapes

    # ---- Dickson frozen benchmark fingerprints (XOR of j)
    c, _ = module_constants("dickson-ladders/score.py")
    shapes_v, at = c["SHAPES"]
    fx["dickson_fingerprints"] = [
        {"label": s[0], "n": s[1], "q2": s[2], "j0": _jsonable(s[3]),
         "span": _jsonable(s[4]), "count": s[5], "xor": _jsonable(s[6]),
         "at": at} for s in shapes_v]

    # ---- AP frozen benchmark fingerprint (XOR of offsets from base)
    c, _ = module_constants("primorial-ap/score.py",
                            env={"W0": 30, "Q2_DEFAULT": 1 << 16})
    fx["ap_fingerprints"] = [{
        "label": "SCORE", "n": c["BENCH_N"][0], "q2": c["BENCH_Q2"][0],
        "base": _jsonable(c["BENCH_BASE"][0]),
        "span": _jsonable(c["BENCH_SPAN"][0]),
        "launch_u": c["BENCH_LAUNCH_U"][0],
        "count": c["FP_COUNT"][0], "xor": _jsonable(c["FP_CHECKSUM"][0]),
        "at": c["FP_COUNT"][1]}]

    # ---- AP measured gate windows (count in the trailing comment)
    src = (REF / "primorial-ap/ap_gpu.py").read_text().splitlines()
    env = {"W0": 30, "P_CEIL": 10**26,
           "_down": lambda p: int(p) - int(p) % 30}
    wins = []
    for i, line in enumerate(src, 1):
        m = re.match(r"^(_W_[A-Z]+)\s*=\s*(\(.*\))\s*#\s*(\d+)\s*$", line)
        if m:
            val = _eval(ast.parse(m.group(2), mode="eval").body, env)
            wins.append({"name": m.group(1), "n": val[0], "q2": val[1],
                         "base": _jsonable(val[2]), "span": _jsonable(val[3]),
                         "count": int(m.group(3)),
                         "at": f"primorial-ap/ap_gpu.py:{i}"})
    fx["ap_gate_windows"] = wins

    # ---- literature tables and floors
    for fam, rel, names in (
            ("euler", "euler-prime-runs/euler_reference.py",
             ("KNOWN", "A21_UPPER", "STATED_FRONTIER")),
            ("dickson", "dickson-ladders/ladder_reference.py",
             ("KNOWN", "PUBLISHED_BOUNDS", "K_FLOOR")),
            ("ap", "primorial-ap/ap_reference.py",
             ("KNOWN", "P_FLOOR", "W0", "W0_RESIDUES"))):
        c, _ = module_constants(rel, names=names)
        fx[fam + "_reference"] = {k: {"value": _jsonable(v), "at": at}
                                  for k, (v, at) in c.items()}

    # ---- engine constants the record semantics depend on
    consts = {}
    for rel, names in (("euler-prime-runs/euler_search.py",
                        ("Q1", "Q2", "P_FLOOR", "P_CEIL", "M_WHEEL",
                         "M_WHEEL_29", "M_WHEEL_31", "M_WHEEL_37")),
                       ("dickson-ladders/ladder_search.py",
                        ("Q2_DEFAULT", "J_CEIL")),
                       ("dickson-ladders/launch.py", ("SPRP_EXACT_FROM",)),
                       ("primorial-ap/ap_search.py",
                        ("Q2_DEFAULT", "P_CEIL", "DEPTH_FLOOR",
                         "DEPTH_EXACT_FROM"))):
        c, _ = module_constants(rel, names=names)
        for k, (v, at) in c.items():
            consts[f"{rel.split('/')[0]}.{k}"] = {"value": _jsonable(v),
                                                  "at": at}
    fx["constants"] = consts

    # ---- discovery ledgers (exact values + the factor that breaks each run)
    ev = {}
    e = json.loads((REF / "euler-prime-runs/evidence/euler_discoveries.json")
                   .read_text())
    ev["euler"] = [{"p": str(r["p"]), "run": r["run"],
                    "breaker_x": r.get("breaker_x"),
                    "breaker_factor": r.get("breaker_factor")} for r in e]
    d = json.loads((REF / "dickson-ladders/evidence/ladder_discoveries.json")
                   .read_text())
    ev["dickson"] = [{"k": str(r["k"]), "run": r["run"],
                      "breaker_m": r.get("breaker_m"),
                      "breaker_factor": r.get("breaker_factor")} for r in d]
    a = json.loads((REF / "primorial-ap/evidence/ap_discoveries.json")
                   .read_text())
    ev["ap"] = [{"p": str(r["p"]), "n": r["n"], "depth": r["depth"],
                 "difference": str(r["difference"])} for r in a]
    # the verification outputs recorded with each discovery: the values and
    # breakers, Dickson's factorizations of k and BLS75 witnesses, AP's three
    # legs and per-value proofs
    ev["euler_full"] = [_jsonable({k: r[k] for k in ("p", "run", "values_prime_x",
                                                 "breaker_x", "breaker",
                                                 "breaker_factor")})
                        for r in e]
    ev["dickson_full"] = [_jsonable({k: r[k] for k in (
        "k", "run", "values_prime_m", "k_factorization", "breaker_m",
        "breaker_factor", "certificates_bls75")}) for r in d]
    ev["ap_full"] = [_jsonable({k: r[k] for k in (
        "p", "n", "depth", "difference", "values", "legs", "agree", "proofs")
        if k in r}) for r in a]
    ev["_at"] = {"euler": "euler-prime-runs/evidence/euler_discoveries.json",
                 "dickson": "dickson-ladders/evidence/ladder_discoveries.json",
                 "ap": "primorial-ap/evidence/ap_discoveries.json"}
    fx["evidence"] = ev

    OUT.write_text(json.dumps(fx, indent=1, sort_keys=True) + "\n")
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
