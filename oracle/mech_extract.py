"""Test infrastructure (oracle): the generator of two data tables the product is
built from, run by hand where the reference tree is present; its outputs are
committed under unicorn_amd/mech/.

1. The San Diego 52-species mechanism of UNICORN's chemistry step (``sd52.json``),
   which the HIP kernels and the CPU oracle are built from.
2. The transport-property tables of the main program (``transport.json``): the
   Lennard-Jones diameters CDIA and well depths TLJ (main:97-108 DATA), the
   collision-integral fits CI11 (main:109-113 DATA) and the conductivity/viscosity
   integral fits CISP of /CB08/ (src/lib/DATA.f), read by main:829-1005.

The reference encodes the mechanism partly as data (``src/lib/DATA.f``: Arrhenius
and fall-off parameters, NASA-7 polynomials) and partly as straight-line code:

* ``src/lib/ckrates.f:117-641``  reverse rates (which species enter K_p, and the
  RUT = 82.06 T power)
* ``src/lib/ckrates.f:645-731``  third-body efficiency sums TB001..TB010
* ``src/lib/ckrates.f:740-759``  third-body group of each fall-off reaction
* ``src/lib/ckrates.f:796-1333`` RC assembly (molecularity, third body, 1/W_k)
* ``src/lib/solvesp.f:87-2806``  the point-implicit Gauss-Seidel schedule
* ``src/lib/solveh.f:117-1131``  RA products and PRD production sums

This script parses those statements (fixed-form Fortran, read as text), derives
each reaction's reactants and products, cross-checks every file against the
others and against element balance, and symbolically executes one SOLVESP
cell-sweep to record which stored reaction products are *stale* when they are
read (the reference's Gauss-Seidel quirks).  It runs only where the reference
tree is present; its output is committed.

Numeric constants (A, b, E/R, fall-off parameters, NASA-7 tables) are NOT taken
from here: they are dumped by the flang-compiled reference itself (oracle/ref
driver) so that float32-widened literals are exact.

Usage:  python oracle/mech_extract.py --write   (both tables; needs oracle/_ref/ref_driver)
"""
from __future__ import annotations

import json
import re
import sys
from collections import Counter, OrderedDict
from pathlib import Path

SUFFIX = ([f"A{i}" for i in range(10)] + [f"A{c}" for c in "ABCDEFGHIJ"]
          + [f"B{i}" for i in range(10)] + [f"B{c}" for c in "ABCDEFGHIJ"]
          + [f"C{i}" for i in range(10)] + ["CA", "CB"])
assert len(SUFFIX) == 52
SFX2K = {s: k + 1 for k, s in enumerate(SUFFIX)}          # 'A0' -> 1 (H2)

NAMES = ["H2", "O2", "H", "O", "OH", "H2O", "HO2", "H2O2", "CO", "CO2", "HCO",
         "CH2O", "CH4", "CH3", "T-CH2", "S-CH2", "C2H4", "CH3O", "C2H5", "C2H6",
         "CH", "C2H2", "C2H3", "CH2CHO", "C2H4O", "CH2CO", "HCCO", "C2H", "CH2OH",
         "CH3OH", "C2H5OH", "CH3CHO", "CH3CHOH", "CH2CH2OH", "CH3CO", "CH3CH2O",
         "C3H4", "C3H3", "C3H5", "C3H6", "C3H8", "I-C3H7", "N-C3H7", "C4H6", "CHO",
         "C5H8", "C7H16", "C4H8", "C5H10", "AR", "HE", "N2"]   # README.md:20-31


def formula(name: str) -> dict:
    """Element counts from a species name (README.md:20-31); CHO is HCO."""
    core = name.split("-", 1)[1] if name[:2] in ("T-", "S-", "I-", "N-") else name
    if core in ("AR", "HE"):
        return {core: 1}
    out: Counter = Counter()
    for el, n in re.findall(r"([A-Z][a-z]?)(\d*)", core.replace("AR", "Ar").replace("HE", "He")):
        out[el] += int(n) if n else 1
    return dict(out)


# --------------------------------------------------------------------------- #
# fixed-form reader
# --------------------------------------------------------------------------- #
def statements(path: Path):
    """Return [(first_line_no, label, code)] with continuations joined,
    comments dropped, blanks removed and text upper-cased."""
    out = []
    for no, raw in enumerate(path.read_text().splitlines(), 1):
        if not raw.strip() or raw[0] in "Cc*!":
            continue
        line = raw[:72].ljust(7)
        label, cont, code = line[:5].strip(), line[5], line[6:]
        if cont not in (" ", "0") and out:
            out[-1][2] += code
        else:
            out.append([no, label, code])
    return [(no, lab, re.sub(r"\s+", "", code).upper()) for no, lab, code in out]


# --------------------------------------------------------------------------- #
# ckrates.f
# --------------------------------------------------------------------------- #
RE_REV = re.compile(r"^SPRF\((\d+)\)=AKF\((\d+)\)\*SPRF\((\d+)\)((?:[*/]RUT)*)/\((.*)\)$")
RE_RC = re.compile(r"^RC\(IA,(\d+)\)=C(1ST|2ND|3RD)\*RHO([123])\*(?:(TB\w+)\*)?SPRF\((\d+)\)"
                   r"((?:\*F\w\w\(J,I\))*)((?:/WM\(\d+\))*)$")
RE_TB = re.compile(r"^(TB\w+)=(.*)$")
RE_DMOLE = re.compile(r"^IF\(IR\.EQ\.(\d+)\)DMOLE=(TB\w+)\*DMOLEA$")


def parse_gsp_quotient(expr: str):
    """'GSP(5)*GSP(4)/GSP(3)/GSP(2)' -> ([5,4], [3,2]) (numerator, denominator)."""
    num, den = [], []
    for op, k in re.findall(r"([*/]?)GSP\((\d+)\)", expr):
        (den if op == "/" else num).append(int(k))
    ops = re.findall(r"([*/]?)GSP\(\d+\)", expr)
    # all factors of the numerator are written before the first division
    assert ops == [""] + ["*"] * (len(num) - 1) + ["/"] * len(den), expr
    check = re.sub(r"[*/]?GSP\(\d+\)", "", expr)
    assert check == "", f"unparsed gsp quotient {expr!r}"
    return num, den


def parse_ckrates(path: Path):
    st = statements(path)
    rev, rc, tb, dmole = {}, {}, OrderedDict(), {}
    for no, _lab, s in st:
        m = RE_REV.match(s)
        if m:
            slot, akf, fwd, rut, q = int(m[1]), int(m[2]), int(m[3]), m[4], m[5]
            assert akf == slot and fwd == slot - 1 and slot % 2 == 0, s
            dn = rut.count("*RUT") - rut.count("/RUT")
            num, den = parse_gsp_quotient(q)
            rev[slot] = {"line": no, "dn": dn, "gsp_num": num, "gsp_den": den,
                         "rut_text": rut}
            continue
        m = RE_RC.match(s)
        if m:
            slot = int(m[1])
            assert int(m[5]) == slot, s
            extra = [SFX2K[x] for x in re.findall(r"\*F(\w\w)\(J,I\)", m[6])]
            wm = [int(x) for x in re.findall(r"/WM\((\d+)\)", m[7])]
            rc[slot] = {"line": no, "order": {"1ST": 1, "2ND": 2, "3RD": 3}[m[2]],
                        "rho": int(m[3]), "tb": m[4], "extra_old": extra, "wm": wm}
            continue
        m = RE_DMOLE.match(s)
        if m:
            dmole[int(m[1])] = m[2]
            continue
        m = RE_TB.match(s)
        if m and (m[1] == "TBODY" or re.fullmatch(r"TB\d{3}", m[1])):
            name, rhs = m[1], m[2]
            terms = []
            body = rhs
            if name == "TBODY" and body.startswith("TBODY"):
                body = body[len("TBODY"):]
            elif name != "TBODY":
                assert body.startswith("TBODY+") or body.startswith("TBODY-"), s
                body = body[len("TBODY"):]
            if not body.startswith(("+", "-")):
                body = "+" + body
            for sign, coef, sfx, w in re.findall(
                    r"([+-])(?:([0-9.]+)\*)?F(\w\w)\(J,I\)/WM\((\d+)\)", body):
                k = SFX2K[sfx]
                assert int(w) == k, s
                terms.append({"sign": sign, "coef": coef or None, "species": k})
            assert re.sub(r"([+-])(?:([0-9.]+)\*)?F(\w\w)\(J,I\)/WM\((\d+)\)", "", body) == "", s
            tb.setdefault(name, {"line": no, "terms": []})["terms"].extend(terms)
    return rev, rc, tb, dmole


# --------------------------------------------------------------------------- #
# solveh.f
# --------------------------------------------------------------------------- #
RE_RA = re.compile(r"^RA\((\d+)\)=RC\(IA,(\d+)\)((?:\*SSP\(IA,\d+\))*)$")
RE_PRD = re.compile(r"^PRD\((\d+)\)=WM\((\d+)\)\*HFO\((\d+)\)\*\((.*)\)$")


def parse_solveh(path: Path):
    ra, prd = {}, {}
    for no, _lab, s in statements(path):
        m = RE_RA.match(s)
        if m:
            slot = int(m[1])
            assert int(m[2]) == slot
            ra[slot] = {"line": no, "ssp": [int(x) for x in re.findall(r"SSP\(IA,(\d+)\)", m[3])]}
            continue
        m = RE_PRD.match(s)
        if m:
            k = int(m[1])
            assert int(m[2]) == k and int(m[3]) == k
            terms = []
            body = m[4]
            pat = r"([+-])(?:([0-9.]+)\*)?(RA|SRA)\((\d+)\)"
            for sign, coef, kind, idx in re.findall(pat, body):
                terms.append({"sign": sign, "coef": coef or None, "kind": kind, "idx": int(idx)})
            assert re.sub(pat, "", body) == "", body
            prd[k] = {"line": no, "terms": terms}
    return ra, prd


# --------------------------------------------------------------------------- #
# solvesp.f : symbolic execution of one cell-sweep
# --------------------------------------------------------------------------- #
RE_A_PROD = re.compile(r"^A(\d{4})=((?:S\w\w\(J,I\)\*)*)RC\(J,I,(\d+)\)$")
RE_A_ZERO = re.compile(r"^A(\d{4})=0\.0$")
RE_A_FNEW = re.compile(r"^A(\d{4})=FNEW\*A(\d{4})$")
RE_RHS1A = re.compile(r"^RHS1A=(.*)$")
RE_RHS1B = re.compile(r"^RHS1B=(.*)$")
RE_RHS1 = re.compile(r"^RHS1=RHO\(J,I\)\*F(\w\w)\(J,I\)\+DT\*WM\((\d+)\)\*\(RHS1A\+RHS1B\)$")
RE_RHS2 = re.compile(r"^RHS2=-DT\*WM\((\d+)\)\*\((.*)\)$")
RE_SET = re.compile(r"^S(\w\w)\(J,I\)=FNEW$")
RE_TERM = re.compile(r"([+-])(?:([0-9.]+)\*)?(?:A(\d{4})|SOUR\(J,I,(\d)\)(?:\*S(\w\w)\(J,I\))?)")


def parse_sum(body: str):
    terms = []
    for sign, coef, a, sour, sfx in RE_TERM.findall(body):
        if a:
            terms.append({"sign": sign, "coef": coef or None, "slot": int(a)})
        else:
            terms.append({"sign": sign, "coef": coef or None, "sour": int(sour),
                          "times": SFX2K[sfx] if sfx else None})
    assert RE_TERM.sub("", body) == "", f"unparsed sum {body!r}"
    return terms


def symbolic_sweep(path: Path):
    """Execute the SOLVESP cell body once, symbolically.

    Every A-variable holds a sorted tuple of factors ('RC', r) / ('S', k, v), v
    being the version of species k's iterate (0 = value at the start of this
    cell-sweep, 1 = value written by this sweep's block for k).  Returns the
    initial-product slots and the block list; every term read carries the
    factor tuple it actually had at that point (``value``).
    """
    st = statements(path)
    start = next(i for i, (no, lab, s) in enumerate(st) if s == "IF(ISKIP(J,I).GE.3)GOTO305")
    end = next(i for i, (no, lab, s) in enumerate(st) if lab == "305")
    body = st[start + 1:end]

    version = {k: 0 for k in range(1, 53)}
    A: dict = {}
    initial: list = []
    blocks: list = []
    pending: list = []
    cur = None
    started = False

    def new_block(kind):
        nonlocal cur, pending
        cur = {"loss_form": pending, "post": [], "kind": kind}
        pending = []
        blocks.append(cur)

    def val(t):
        return dict(t, value=A.get(t["slot"], "UNSET")) if "slot" in t else t

    for no, _lab, s in body:
        m = RE_A_PROD.match(s)
        if m:
            slot, sfx, r = int(m[1]), m[2], int(m[3])
            assert slot == r, s
            ks = [SFX2K[x] for x in re.findall(r"S(\w\w)\(J,I\)", sfx)]
            A[slot] = tuple(sorted([("RC", r)] + [("S", k, version[k]) for k in ks]))
            if started:
                pending.append({"slot": slot, "s_factors": ks, "line": no})
            else:
                initial.append({"slot": slot, "s_factors": ks})
            continue
        m = RE_A_ZERO.match(s)
        if m:
            assert not started
            A[int(m[1])] = ("ZERO",)
            initial.append({"slot": int(m[1]), "zero": True})
            continue
        m = RE_A_FNEW.match(s)
        if m:
            a, b = int(m[1]), int(m[2])
            assert a == b and cur is not None and isinstance(cur.get("k"), int)
            A[a] = tuple(sorted(list(A[a]) + [("S", cur["k"], 1)]))
            cur["post"].append(a)
            continue
        if s.startswith("RHS1A="):
            started = True
            new_block("species")
            cur["rhs1a"] = [val(t) for t in parse_sum(RE_RHS1A.match(s)[1])]
            continue
        if s.startswith("RHS1B="):
            if cur is None or "set" in cur:
                # species 50/51: no RHS1A statement; the previous block's RHS1A is reused
                new_block("species_reuse_rhs1a")
                cur["rhs1a"] = "REUSE_PREVIOUS"
            b = RE_RHS1B.match(s)[1]
            cur["rhs1b"] = [] if b == "0.0" else [val(t) for t in parse_sum(b)]
            continue
        m = RE_RHS1.match(s)
        if m:
            cur["k"] = SFX2K[m[1]]
            assert int(m[2]) == cur["k"], s
            continue
        if s.startswith("RHS2="):
            if s == "RHS2=0.0":
                cur["rhs2"] = []
            else:
                m = RE_RHS2.match(s)
                if m:
                    assert int(m[1]) == cur["k"]
                    cur["rhs2"] = [val(t) for t in parse_sum(m[2])]
                else:
                    cur["rhs2"] = ["SOOT", s]
            continue
        m = RE_SET.match(s)
        if m:
            k = SFX2K[m[1]]
            assert k == cur["k"], (s, cur["k"])
            version[k] = 1
            cur["set"] = True
            continue
        if s.startswith("RHS1=RHO(J,I)*STMF") or s.startswith("RHS1=RHO(J,I)*STND"):
            new_block("soot")
            cur["k"] = "SST1" if "STMF" in s else "SST2"
            cur["rhs1"] = s
            continue
        if s.startswith("SST1(J,I)=FNEW") or s.startswith("SST2(J,I)=FNEW"):
            cur["set"] = True
            continue
    assert not pending, pending
    return initial, blocks, body


def main(ref_root="/root/reference", out=None):
    ref = Path(ref_root) / "src" / "lib"
    rev, rc, tb, dmole = parse_ckrates(ref / "ckrates.f")
    ra, prd = parse_solveh(ref / "solveh.f")
    initial, blocks, body = symbolic_sweep(ref / "solvesp.f")
    return rev, rc, tb, dmole, ra, prd, initial, blocks




# --------------------------------------------------------------------------- #
# assembling the mechanism and cross-checking the three files
# --------------------------------------------------------------------------- #
class MechError(AssertionError):
    pass


def check(cond, msg):
    if not cond:
        raise MechError(msg)


def build(ref_root="/root/reference"):
    ref = Path(ref_root) / "src" / "lib"
    rev, rc, tb, dmole = parse_ckrates(ref / "ckrates.f")
    ra, prd = parse_solveh(ref / "solveh.f")
    initial, blocks, _body = symbolic_sweep(ref / "solvesp.f")

    # reactants of every written slot, from SOLVEH's RA products
    reac = {s: sorted(v["ssp"]) for s, v in ra.items()}
    check(set(reac) == set(rc), "RA and RC slot sets differ")
    for s, v in rc.items():
        check(sorted(v["wm"]) == sorted(reac[s] + v["extra_old"]),
              f"slot {s}: RC 1/W list {v['wm']} != RA reactants {reac[s]} + {v['extra_old']}")
        n = len(v["wm"]) + (1 if v["tb"] else 0)
        check(v["order"] == n and v["rho"] == n, f"slot {s}: molecularity {v['order']} != {n}")

    # products: PRD sums; lumped terms carry coefficient 1.000
    prod_of = {s: [] for s in range(1, 545)}
    loss_of = {s: [] for s in range(1, 545)}
    for k, v in prd.items():
        for t in v["terms"]:
            if t["kind"] != "RA":
                continue
            check(t["coef"] in (None, "1.000"), f"PRD({k}) coefficient {t['coef']}")
            (prod_of if t["sign"] == "+" else loss_of)[t["idx"]].append(k)
    for s in range(1, 545):
        if s in reac:
            check(sorted(loss_of[s]) == reac[s],
                  f"slot {s}: PRD loss species {sorted(loss_of[s])} != reactants {reac[s]}")
    prods = {s: sorted(prod_of[s]) for s in range(1, 545)}

    # reverse forms: K_p numerator = products of the forward slot, denominator = reactants
    for s, v in rev.items():
        f = s - 1
        spect = rc[f]["extra_old"]          # C3H8 collider of 459/460 appears on both sides of K_p
        check(sorted(v["gsp_num"]) == sorted(prods[f] + spect),
              f"rev {s}: GSP numerator {v['gsp_num']} != products of {f} {prods[f]}")
        check(sorted(v["gsp_den"]) == sorted(reac[f] + spect),
              f"rev {s}: GSP denominator {v['gsp_den']} != reactants of {f}")
        check(v["dn"] == len(v["gsp_num"]) - len(v["gsp_den"]), f"rev {s}: RUT power {v['dn']}")
        check(rc[s]["extra_old"] == spect, f"rev {s}: spectator mismatch")
        check(prods[s] == reac[f] and reac.get(s) == prods[f],
              f"rev {s}: reverse slot stoichiometry is not the mirror of {f}")

    # element balance (lumped C7H16 steps are reported, not required)
    comp = [formula(n) for n in NAMES]
    unbalanced = []
    for s in range(1, 545, 2):
        lhs, rhs = Counter(), Counter()
        for k in reac.get(s, []):
            lhs.update(comp[k - 1])
        for k in prods[s]:
            rhs.update(comp[k - 1])
        if lhs != rhs:
            unbalanced.append(s)

    # third-body groups
    tb_of = {s: v["tb"] for s, v in rc.items() if v["tb"]}
    # ------------------------------------------------------------------ sweep
    species_blocks = [b for b in blocks if b["kind"] != "soot"]
    check([b["k"] for b in species_blocks] == list(range(1, 52)), "species block order is not 1..51")
    updated = set()
    stale = []          # production terms whose stored product is not fully current
    for b in species_blocks:
        k = b["k"]
        if b["kind"] == "species":
            cur_ver = {j: (1 if j in updated else 0) for j in range(1, 53)}
            # loss terms: RC_s * prod_{j in reac(s) minus one k} S_j(current)
            exp_loss = Counter()
            for s, rs in reac.items():
                exp_loss[s] += rs.count(k)
            got_loss = Counter()
            for t in b["rhs2"]:
                if "slot" not in t:
                    continue
                check(t["sign"] == "+" and t["coef"] is None, f"block {k}: loss term {t}")
                s = t["slot"]
                got_loss[s] += 1
                others = list(reac[s])
                others.remove(k)
                want = tuple(sorted([("RC", s)] + [("S", j, cur_ver[j]) for j in others]))
                check(t["value"] == want, f"block {k}: loss A{s:04d} = {t['value']} expected {want}")
            check(+got_loss == +exp_loss, f"block {k}: loss set {sorted((+got_loss).items())} != {sorted((+exp_loss).items())}")
            # production terms: RC_s * prod_{j in reac(s)} S_j(current)
            exp_prod = Counter()
            for s in range(1, 545):
                exp_prod[s] += prods[s].count(k)
            got_prod = Counter()
            for t in list(b["rhs1a"]) + list(b["rhs1b"]):
                if "slot" not in t:
                    continue
                check(t["sign"] == "+" and t["coef"] in (None, "1.000"), f"block {k}: prod term {t}")
                s = t["slot"]
                got_prod[s] += 1
                if s not in reac:
                    check(t["value"] == ("ZERO",), f"block {k}: unwritten slot {s} read as {t['value']}")
                    continue
                want = tuple(sorted([("RC", s)] + [("S", j, cur_ver[j]) for j in reac[s]]))
                if t["value"] != want:
                    stale.append({"block": k, "slot": s, "got": t["value"], "want": want})
            check(+got_prod == +exp_prod, f"block {k}: production set differs: "
                  f"extra {sorted((got_prod - exp_prod).items())} missing {sorted((exp_prod - got_prod).items())}")
        updated.add(k)
    return {"rev": rev, "rc": rc, "tb": tb, "dmole": dmole, "ra": ra, "prd": prd,
            "initial": initial, "blocks": blocks, "reac": reac, "prods": prods,
            "unbalanced": unbalanced, "stale": stale, "tb_of": tb_of}


if __name__ == "__main__" and len(sys.argv) > 1 and sys.argv[1] == "--check":
    m = build()
    print("unbalanced:", m["unbalanced"])
    print("stale:", len(m["stale"]))
    for x in m["stale"]:
        print(x)


# --------------------------------------------------------------------------- #
# numeric constants dumped by the flang-built reference (oracle ref_driver mech)
# --------------------------------------------------------------------------- #
def read_mech_dump(path):
    import numpy as np
    raw = Path(path).read_bytes()
    off = 0

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=off)
        off += a.nbytes
        return a

    lrev, nrlind = take("<i4", 2)
    itbef = take("<i4", 544)
    arr = {name: take("<f8", 544) for name in ("ALF", "AKF", "EXAF", "ALLOW", "AKLOW", "EALOW")}
    troe = take("<f8", 4 * 544).reshape(4, 544)
    tpol = take("<f8", 2)
    polsp = take("<f8", 14 * 52).reshape(52, 14)
    cisp = take("<f8", 12 * 52).reshape(52, 12)
    cb06 = take("<f8", 20)
    cb02 = take("<f8", 4 + 52)
    hfo = take("<f8", 52)
    assert off == len(raw), (off, len(raw))
    return dict(lrev=int(lrev), nrlind=int(nrlind), itbef=itbef, troe=troe, tpol=tpol,
                polsp=polsp, cisp=cisp, cb06=cb06, wm=cb02[4:], hfo=hfo, **arr)


def h(x):
    return float(x).hex()


def term_json(t):
    out = {"sign": t["sign"]}
    if t.get("coef"):
        out["coef"] = t["coef"]
    if "slot" in t:
        out["slot"] = t["slot"]
    else:
        out["sour"] = t["sour"]
        if t.get("times"):
            out["times"] = t["times"]
    return out


def to_table(m, dump):
    import numpy as np
    species = []
    for k in range(52):
        species.append({"k": k + 1, "name": NAMES[k], "formula": formula(NAMES[k]),
                        "nasa7": [h(v) for v in dump["polsp"][k]]})
    reactions = []
    falloff_tb = {s: m["dmole"].get(s, "TBODY") for s in range(1, 545, 2) if dump["AKLOW"][s - 1] != 0.0}
    for f in range(1, 545, 2):
        r = f + 1
        rev = m["rev"].get(r)
        ent = {
            "f": f,
            "reac": m["reac"][f], "prod": m["prods"][f],
            "spectator_old": m["rc"][f]["extra_old"],
            "A": h(dump["AKF"][f - 1]), "b": h(dump["ALF"][f - 1]), "E_R": h(dump["EXAF"][f - 1]),
            "rc_fwd": {"order": m["rc"][f]["order"], "tb": m["rc"][f]["tb"], "wm": m["rc"][f]["wm"]},
            "rev": None,
            "falloff": None,
        }
        if r in m["rc"]:
            ent["rc_rev"] = {"order": m["rc"][r]["order"], "tb": m["rc"][r]["tb"], "wm": m["rc"][r]["wm"]}
            ent["rev_reac"] = m["reac"][r]
        else:
            ent["rc_rev"] = None
        if rev is not None:
            ent["rev"] = {"dn": rev["dn"], "gsp_num": rev["gsp_num"], "gsp_den": rev["gsp_den"],
                          "form": rev["rut_text"] or "", "line": rev["line"]}
            check(dump["AKF"][r - 1] == 1.0, f"AKF({r}) != 1")
        if f in falloff_tb:
            ent["falloff"] = {"tb": falloff_tb[f], "A0": h(dump["AKLOW"][f - 1]),
                              "b0": h(dump["ALLOW"][f - 1]), "E0_R": h(dump["EALOW"][f - 1]),
                              "troe": [h(dump["troe"][t][f - 1]) for t in range(4)]}
        else:
            check(all(dump["troe"][t][f - 1] == 0 for t in range(4)), f"troe set on non-falloff {f}")
        reactions.append(ent)
    tbg = {name: [{"sign": t["sign"], "coef": t["coef"], "species": t["species"]} for t in v["terms"]]
           for name, v in m["tb"].items()}
    blocks = []
    for b in m["blocks"]:
        if b["kind"] == "soot":
            blocks.append({"k": b["k"], "kind": "soot"})
            continue
        blocks.append({
            "k": b["k"], "kind": b["kind"],
            "loss_form": [{"slot": x["slot"], "s_factors": x["s_factors"]} for x in b["loss_form"]],
            "rhs1a": None if b["rhs1a"] == "REUSE_PREVIOUS" else [term_json(t) for t in b["rhs1a"]],
            "rhs1b": [term_json(t) for t in b["rhs1b"]],
            "rhs2": [term_json(t) for t in b["rhs2"]],
            "post": b["post"],
        })
    stale = [{"block": x["block"], "slot": x["slot"]} for x in m["stale"]]
    prd = {}
    for k, v in sorted(m["prd"].items()):
        prd[str(k)] = [{"sign": t["sign"], "coef": t["coef"], "kind": t["kind"], "idx": t["idx"]}
                       for t in v["terms"]]
    wm_ok = np.allclose(dump["wm"], np.array([w for w in WSP]) / 1000.0 / dump["cb06"][13], rtol=0, atol=0)
    return {
        "about": "San Diego 52-species / 544-slot mechanism of UNICORN's chemistry step "
                 "(src/lib/DATA.f, src/lib/ckrates.f, src/lib/solvesp.f, src/lib/solveh.f); "
                 "numbers are those of the flang-compiled reference after main:364-413.",
        "nspecies": 52, "nslots": 544,
        "tpol": [h(dump["tpol"][0]), h(dump["tpol"][1])],
        "species": species,
        "reactions": reactions,
        "written_rc_slots": sorted(m["rc"]),
        "unassigned_reverse": sorted(set(range(2, 545, 2)) & set(m["rc"]) - set(m["rev"])),
        "tb_groups": tbg,
        "sweep": {"initial": m["initial"], "blocks": blocks, "stale_doubled": stale},
        "prd": prd,
        "ra_factors": {str(k): v["ssp"] for k, v in sorted(m["ra"].items())},
        "wm_check": bool(wm_ok),
    }


WSP = [2.016, 32.00, 1.008, 16.00, 17.01, 18.02, 33.01, 34.01, 28.01, 44.01, 29.02, 30.03, 16.04,
       15.04, 14.03, 14.03, 28.05, 31.03, 29.06, 30.07, 13.02, 26.04, 27.05, 43.05, 44.05, 42.04,
       41.03, 25.03, 31.03, 32.04, 46.07, 44.05, 45.06, 45.06, 43.05, 45.06, 40.07, 39.06, 41.07,
       42.08, 44.10, 43.09, 43.09, 54.09, 29.02, 68.12, 100.2, 56.11, 70.14, 39.95, 4.003, 28.01]


def main_data(name: str, ref_root="/root/reference") -> list:
    """The values of one DATA statement of the main program (all REAL*8 literals with
    a D exponent, read exactly)."""
    prog = Path(ref_root) / "src" / "unicornd-mild-sd-5.f"
    for _, _, code in statements(prog):
        m = re.match(r"^DATA%s/(.*)/$" % re.escape(name), code)
        if m:
            vals = [v.replace("D", "E") for v in m.group(1).split(",")]
            return [float(v) for v in vals]
    raise MechError(f"DATA {name} not found")


def transport_table(dump, ref_root="/root/reference") -> dict:
    """CDIA (52), TLJ (52), CI11 (4x5, the DATA's column-major order turned into rows
    K=1..4) and CISP (52x12, from the flang build's /CB08/), all as hex floats."""
    cdia, tlj, ci = main_data("CDIA", ref_root), main_data("TLJ", ref_root), main_data("CI11", ref_root)
    check(len(cdia) == 52 and len(tlj) == 52 and len(ci) == 20, "transport DATA sizes")
    ci11 = [[ci[k + 4 * m] for m in range(5)] for k in range(4)]      # CI11(K, M), K fastest in DATA
    return {"about": "transport-property tables of UNICORN's main program: CDIA, TLJ "
                     "(src/unicornd-mild-sd-5.f DATA), CI11 (K=1..4 rows, M=1..5 columns), "
                     "CISP(12,52) of /CB08/ (src/lib/DATA.f) as the flang build holds it",
            "cdia": [h(v) for v in cdia], "tlj": [h(v) for v in tlj],
            "ci11": [[h(v) for v in row] for row in ci11],
            "cisp": [[h(v) for v in dump["cisp"][k]] for k in range(52)]}


if __name__ == "__main__" and len(sys.argv) > 1 and sys.argv[1] == "--write":
    import subprocess
    import tempfile
    here = Path(__file__).resolve().parent
    drv = here / "_ref" / "ref_driver"
    mech = here.parent / "unicorn_amd" / "mech"
    with tempfile.TemporaryDirectory() as td:
        subprocess.run([str(drv), "mech", f"{td}/mech.bin"], check=True)
        dump = read_mech_dump(f"{td}/mech.bin")
    table = to_table(build(), dump)
    out = mech / "sd52.json"
    out.write_text(json.dumps(table, separators=(",", ":")))
    print("wrote", out, "reactions", len(table["reactions"]), "wm_check", table["wm_check"])
    out = mech / "transport.json"
    out.write_text(json.dumps(transport_table(dump), indent=0))
    print("wrote", out)
