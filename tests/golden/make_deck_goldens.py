"""Known answers for the deck reader (ompVC/parameters.c:20-292): the
reference program's own echo of the values it read (parameters.c:142-230)
from each deck shipped with ompVC.

Run in the authoring container only (needs /root/reference):
    make -f oracle/ref.mk && python tests/golden/make_deck_goldens.py
For each deck, decks/<name>.entries keeps its NAME = VALUE lines (comment
lines dropped) and decks/<name>.echo.json the name/value pairs the reference
printed after reading the whole deck.  The reference then stops on its
missing data files; only the echo is used.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref", "ompflowVC")
SRC = "/root/reference/ompVC"
# The other decks shipped with ompVC do not give known answers: input.in is a
# template with empty values, and 3Dunstructure.in, unstructure2D.in and
# uppervasc_tracers.in predate the num_threads entry (96 entries), which this
# reference reads shifted by one entry from there on.
DECKS = ["cartesian.in"]
ECHO = re.compile(r"^(?:Reading in parameters\.\.\.)?([A-Za-z_][A-Za-z0-9_.\[\]]*)\s*=\s*(.*?)\s*$")


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    out = os.path.join(HERE, "decks")
    os.makedirs(out, exist_ok=True)
    for name in DECKS:
        text = open(os.path.join(SRC, name)).read()
        entries = [ln.strip() for ln in text.splitlines() if ln.strip() and not ln.lstrip().startswith("#")]
        with tempfile.TemporaryDirectory(prefix="deck_") as work:
            r = subprocess.run([REF, os.path.join(SRC, name)], cwd=work, capture_output=True, text=True, timeout=60)
        echo = []
        for ln in r.stderr.splitlines():
            m = ECHO.match(ln)
            if m:
                echo.append([m.group(1), m.group(2)])
        stem = name[:-3]
        with open(os.path.join(out, stem + ".entries"), "w") as f:
            f.write("\n".join(entries) + "\n")
        with open(os.path.join(out, stem + ".echo.json"), "w") as f:
            json.dump({"deck": "ompVC/" + name, "producer": "ompflowVC built by oracle/ref.mk", "echo": echo}, f,
                      indent=0)
        print(name, len(entries), "entries,", len(echo), "echoed values")


if __name__ == "__main__":
    main()
