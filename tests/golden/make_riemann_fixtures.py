"""Convert the reference's 1D two-material Riemann decks into a JSON fixture.

Reads (text only, nothing is executed) from the reference checkout:
  Tests/RiemannProblems/Test{1..12}/input.st               deck parameters
  Tests/RiemannProblems/Test{1..12}/exact/RiemannSolution.txt  exact solver profile
  Tests/RiemannProblems/Test{1..12}/results/line_0001.txt   full-pipeline line plot
  Tests/RiemannProblems/Test{1..12}/log.out                 per-step t and dt
and writes tests/golden/riemann_problems.json; the same three files of the
two-dimensional explosion deck
  Tests/ExplosionShyue98/{input.st,results/line_0001.txt,log.out}
go to tests/golden/shyue98.json.  Run:
  python tests/golden/make_riemann_fixtures.py /root/reference
"""
import json
import os
import re
import sys


def parse_deck(text):
    """Parse the `under Name { key = value; }` grammar of the reference's
    input decks into nested dicts (parser/Parser.y restated for this subset)."""
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    toks = re.findall(r'"[^"]*"|[A-Za-z_][A-Za-z0-9_]*(?:\[\d+\])?|[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?|[{}=;]', text)
    pos = 0

    def block():
        nonlocal pos
        out = {}
        while pos < len(toks) and toks[pos] != "}":
            t = toks[pos]
            if t == "under":
                name = toks[pos + 1]
                assert toks[pos + 2] == "{", (name, toks[pos:pos + 5])
                pos += 3
                out[name] = block()
                assert toks[pos] == "}"
                pos += 1
            else:
                key = t
                assert toks[pos + 1] == "=", (key, toks[pos:pos + 5])
                val = toks[pos + 2]
                pos += 3
                if pos < len(toks) and toks[pos] == ";":
                    pos += 1
                try:
                    val = float(val)
                except ValueError:
                    val = val.strip('"')
                out[key] = val
        return out

    return block()


def parse_table(path):
    header, rows = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("##"):
                header.append(line.rstrip("\n"))
            elif line.strip():
                rows.append([float(x) for x in line.split()])
    return header, rows


def parse_log(path):
    steps = []
    pat = re.compile(r"Step (\d+): t = ([-+0-9.eE]+), dt = ([-+0-9.eE]+), cfl = ([-+0-9eE]+\.[0-9]+[eE][-+][0-9]+)")
    with open(path, errors="replace") as f:
        for line in f:
            m = pat.search(line)
            if m:
                steps.append([int(m.group(1)), float(m.group(2)), float(m.group(3)), float(m.group(4))])
    return steps


def main(ref):
    base = os.path.join(ref, "Tests", "RiemannProblems")
    out = {}
    for n in range(1, 13):
        d = os.path.join(base, "Test%d" % n)
        with open(os.path.join(d, "input.st")) as f:
            deck = parse_deck(f.read())
        eh, er = parse_table(os.path.join(d, "exact", "RiemannSolution.txt"))
        lh, lr = parse_table(os.path.join(d, "results", "line_0001.txt"))
        out["Test%d" % n] = {
            "deck": deck,
            "exact_header": eh, "exact_rows": er,
            "line_header": lh, "line_rows": lr,
            "log_steps": parse_log(os.path.join(d, "log.out")),
            "source": "Tests/RiemannProblems/Test%d" % n,
        }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "riemann_problems.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", dst)
    d = os.path.join(ref, "Tests", "ExplosionShyue98")
    with open(os.path.join(d, "input.st")) as f:
        deck = parse_deck(f.read())
    lh, lr = parse_table(os.path.join(d, "results", "line_0001.txt"))
    shy = {"deck": deck, "line_header": lh, "line_rows": lr, "log_steps": parse_log(os.path.join(d, "log.out")),
           "source": "Tests/ExplosionShyue98"}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shyue98.json")
    with open(dst, "w") as f:
        json.dump(shy, f, indent=0)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
