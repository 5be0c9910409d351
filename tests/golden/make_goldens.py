"""Write golden fixtures by running the reference ompVC (built from its sources
by oracle/ref.mk into oracle/_ref/) on the cases of cases.py.

Run in the authoring container only (needs /root/reference):
    make -f oracle/ref.mk && python tests/golden/make_goldens.py
Each case directory receives the reference's output files, the .IC file where
the case has one, the files a case lists under "keep" (further .IC files,
the FTLE_ICFile), and manifest.json with input digests.
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import cases  # noqa: E402

REF = os.path.join(cases.REPO, "oracle", "_ref", "ompflowVC")


def main(selected):
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    for name in selected or sorted(cases.CASES):
        c = cases.CASES[name]
        work = tempfile.mkdtemp(prefix="golden_")
        try:
            deck_path, data, out = cases.prepare(name, work)
            r = subprocess.run([REF, deck_path], cwd=work, capture_output=True, text=True)
            if r.returncode:
                sys.exit("%s: reference failed\n%s\n%s" % (name, r.stdout[-2000:], r.stderr[-2000:]))
            fx = cases.fixture_dir(name)
            if os.path.isdir(fx):
                shutil.rmtree(fx)
            os.makedirs(fx)
            produced = sorted(os.listdir(out))
            for f in produced:
                if f.endswith(".info"):
                    continue
                shutil.copy(os.path.join(out, f), fx)
            if c.get("ic"):
                shutil.copy(os.path.join(data, c["ic"]), fx)
            for f in c.get("keep", []):  # further files the reference writes beside the inputs
                shutil.copy(os.path.join(data, f), fx)
            man = dict(case=name, nproc=cases.GOLDEN_NPROC, outputs=[f for f in produced if not f.endswith(".info")],
                       ic=c.get("ic"), keep=c.get("keep", []), inputs=cases.input_digests(data),
                       producer="ompVC (ompVC/*.c built by oracle/ref.mk, gcc -O3 -fopenmp)")
            with open(os.path.join(fx, "manifest.json"), "w") as f:
                json.dump(man, f, indent=1, sort_keys=True)
            print(name, "->", len(man["outputs"]), "outputs", "ic" if c.get("ic") else "")
        finally:
            shutil.rmtree(work)


if __name__ == "__main__":
    main(sys.argv[1:])
