"""Extract the numeric golden vectors held by the reference's unit tests.

Reads the reference Rust test modules as text (nothing is compiled, imported or
run) and writes the literal arrays they compare against to
``reference_vectors.json`` beside this script: for every ``vec![...]`` literal
inside a ``#[test]`` function, the list of numbers (or ``[re, im]`` pairs for
``Complex::new(re, im)``), with the file and line range it came from.

Run in a container that has the reference checkout:
    python tests/golden/extract_reference_vectors.py /root/reference
The JSON is committed; tests never read the reference tree.
"""
import json
import os
import re
import sys

FILES = ["src/flds/field.rs", "src/flds/fft_2d.rs", "src/flds/wave_num.rs"]
NUM = r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?"
CPLX = re.compile(r"Complex::new\(\s*(" + NUM + r")\s*,\s*(" + NUM + r")\s*\)")
FN = re.compile(r"^\s*(?:pub\s+)?fn\s+(\w+)")
LABEL = re.compile(r"([\w.]+)\s*(?::\s*Vec<[^>]*>+)?\s*[:=]\s*vec!\[\s*$|([\w.]+)\s*[:=]\s*vec!\[")


def extract(path, rel):
    lines = open(path).read().split("\n")
    out = []
    fn = None
    in_tests = False
    i = 0
    while i < len(lines):
        line = lines[i]
        if "#[cfg(test)]" in line:
            in_tests = True
        m = FN.match(line)
        if m:
            fn = m.group(1)
        if in_tests and "vec![" in line:
            # collect until the bracket that closes this vec!
            start = i
            depth = 0
            buf = []
            j = i
            col = line.index("vec![")
            seg = line[col + 4:]
            while True:
                for ch in seg:
                    buf.append(ch)
                    if ch == "[":
                        depth += 1
                    elif ch == "]":
                        depth -= 1
                        if depth == 0:
                            break
                if depth == 0:
                    break
                j += 1
                seg = "\n" + lines[j]
            body = "".join(buf)[1:-1]
            lab = re.search(r"([\w.]+)\s*(?::\s*Vec<.*>)?\s*[:=]\s*$", line[:col])
            label = lab.group(1) if lab else "anon"
            if ";" not in body and "zero()" not in body:
                cp = CPLX.findall(body)
                if cp:
                    vals = [[float(a), float(b)] for a, b in cp]
                else:
                    vals = [float(x) for x in re.findall(NUM, body)]
                out.append({"file": rel, "lines": [start + 1, j + 1], "test": fn,
                            "label": label, "values": vals})
            i = j + 1
            continue
        i += 1
    return out


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    vecs = []
    for rel in FILES:
        vecs += extract(os.path.join(root, rel), rel)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.json")
    with open(dst, "w") as f:
        json.dump({"source": "pcrumley/spectral_rs unit tests", "vectors": vecs}, f)
    for v in vecs:
        print(v["file"], v["lines"], v["test"], v["label"], len(v["values"]))


if __name__ == "__main__":
    main()
