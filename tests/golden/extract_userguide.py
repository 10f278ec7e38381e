"""Extract the sample-run tables printed in the reference user guide.

Reads /root/reference/Doc/userguide.ps (a document, not code), pulls the
PostScript text strings of each printed sample run and writes
tests/golden/userguide_runs.json with, per run: eigenvalues, residual norms,
error estimates, cluster/ordinal labels, MATVEC count, estimated gap, INFO,
the 'used shift'/'used droptol' lines of every preconditioner setup and the
suggested SHIFT/DROPTOL.  Run once in the build container:

    python tests/golden/extract_userguide.py
"""
import json
import os
import re

PS = "/root/reference/Doc/userguide.ps"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "userguide_runs.json")

# (name, first line, last line, description) -- line ranges of userguide.ps
RUNS = [
    ("EXAMPLE1", 7560, 7620, "samples/EXAMPLE1.f: DPJD, tridiag(5, i-1, 5) N=1000, "
     "ISEARCH=2 SIGMA=0, NEIG=5, MADSPACE=20, TOL=1e-10, DROPTOL=1e-3"),
    ("EXAMPLE2", 7880, 7960, "samples/EXAMPLE2.f: DPJD, same matrix, ISEARCH=0"),
    ("LISTING5_ADAPTIVE", 9300, 9400, "guide listing 5: DPJDREVCOM, ILU of tridiag(5,0,5), "
     "matvec with tridiag(5,0,5)+0.1*ones, ICNTL(2)=0"),
    ("LISTING5_STATIC", 9580, 9630, "guide listing 6: same with ICNTL(2)=1"),
    ("EXAMPLE9", 12200, 12260, "samples/EXAMPLE9.f: DPJDREVCOM diagonal preconditioner, "
     "tridiag(5, i, 5) N=1000, ISEARCH=0, NEIG=5, MADSPACE=20, TOL=1e-10, ICNTL=0"),
    ("EXAMPLE10", 13100, 13180, "samples/EXAMPLE10.f: as EXAMPLE9 with ICNTL(1)=2, ICNTL(2)=1"),
    ("EXAMPLE11", 14020, 14100, "samples/EXAMPLE11.f: DPJD TOL=1e-3 then ICNTL(2)=-1, "
     "NINIT=NEIG, TOL=1e-10"),
]

NUM = r"[-+]?\d\.\d+E[-+]\d+"


def page_text(lines, a, b):
    chunk = "".join(lines[a - 1:b])
    return " ".join(re.findall(r"\(([^)]*)\)", chunk))


def parse_runs(text):
    out = []
    for block in text.split("**ENTERING PJDREVCOM")[1:]:
        run = {"precond_setups": [], "table": []}
        for m in re.finditer(r"used droptol: (" + NUM + r")", block):
            run["precond_setups"].append({"droptol": float(m.group(1))})
        shifts = [float(m.group(1)) for m in re.finditer(r"used shift:\s*(" + NUM + r")", block)]
        if run["precond_setups"] and len(shifts) == len(run["precond_setups"]):
            for s, sh in zip(run["precond_setups"], shifts):
                s["shift"] = sh
        else:
            run["precond_setups"] = [{"shift": sh} for sh in shifts]
        fills = [float(m.group(1)) for m in re.finditer(r"fill-in factor: ([\d.]+)", block)]
        for s, f in zip(run["precond_setups"], fills):
            s["fill"] = f
        for m in re.finditer(r"(" + NUM + r") (\d\.\dE[-+]\d+) (\d\.\dE[-+]\d+) (-?\d+)\\", block):
            run["table"].append({"eig": float(m.group(1)), "res": m.group(2),
                                 "errest": m.group(3), "no": int(m.group(4))})
        m = re.search(r"Number of MATVEC needed: (\d+) ; Estimated gap: (" + r"[-+]?\d\.\d+E[-+]\d+" +
                      r") ; INFO: (-?\d+)", block)
        if m:
            run["matvec"] = int(m.group(1))
            run["gap"] = m.group(2)
            run["info"] = int(m.group(3))
        m = re.search(r"Suggested SHIFT\s+if restart needed:\s*([-+]?\d\.\d+E[-+]\d+)", block)
        if m:
            run["suggested_shift"] = m.group(1)
        m = re.search(r"Suggested DROPTOL if restart needed:\s*([-+]?\d\.\d+E[-+]\d+)", block)
        if m:
            run["suggested_droptol"] = m.group(1)
        if run["table"]:
            out.append(run)
    return out


def main():
    with open(PS, "r", errors="replace") as f:
        lines = f.readlines()
    data = {"source": "Doc/userguide.ps (JADAMILU user guide, built 2006-10-06)", "runs": {}}
    for name, a, b, desc in RUNS:
        runs = parse_runs(page_text(lines, a, b))
        data["runs"][name] = {"lines": [a, b], "description": desc, "calls": runs}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    for k, v in data["runs"].items():
        print(k, [(c.get("matvec"), len(c["table"]), len(c["precond_setups"])) for c in v["calls"]])


if __name__ == "__main__":
    main()
