"""Writes the golden fixtures under tests/golden/ from the reference's own test data.

Run once in the build container (it reads the reference tree, which is not
present where the tests run):

  python tests/golden/make_golden.py /root/reference

* corpus/synth-*.gz + corpus_sums.json: the reference's committed gzip
  fixtures (tests/corpus/) with their recorded sha256 / sizes
  (tests/corpus/reference_sums.json, required by tests/golden_hash.rs:145-156).
* miri_vectors.json: the raw-DEFLATE known-answer vectors of
  crates/rusty-rapidgzip/tests/miri_edge.rs:18-477, extracted as data
  (name, stream bytes, expected plaintext bytes).
"""
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref: str) -> None:
    src = os.path.join(ref, "tests", "corpus")
    dst = os.path.join(HERE, "corpus")
    os.makedirs(dst, exist_ok=True)
    sums = json.load(open(os.path.join(src, "reference_sums.json")))
    keep = {}
    for name in sorted(os.listdir(src)):
        if name.startswith("synth-") and name.endswith(".gz"):
            shutil.copyfile(os.path.join(src, name), os.path.join(dst, name))
            keep[name] = sums[name]
    json.dump(keep, open(os.path.join(HERE, "corpus_sums.json"), "w"), indent=1, sort_keys=True)

    text = open(os.path.join(ref, "crates", "rusty-rapidgzip", "tests", "miri_edge.rs")).read()
    body = text[text.index("vec![") + 5:]
    vecs = []
    for m in re.finditer(r'\(\s*"(\w+)",\s*vec!\[([^\]]*)\],\s*vec!\[([^\]]*)\],?\s*\)', body):
        def nums(s):
            return [int(x) for x in re.findall(r"\d+", s)]
        vecs.append({"name": m.group(1), "stream": nums(m.group(2)), "expected": nums(m.group(3))})
    json.dump(vecs, open(os.path.join(HERE, "miri_vectors.json"), "w"))
    print(f"{len(keep)} corpus fixtures, {len(vecs)} raw-DEFLATE vectors")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
