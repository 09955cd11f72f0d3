"""Write the golden fixtures under tests/golden/<case>/ from runs of the compiled reference.

Run in the build container, where /root/reference exists:

    make -C oracle ref          # oracle/_ref/swehr_ref and swehr_ref_hooked
    python tests/golden/make_golden.py [case ...]

For every case in cases.CASES it writes the case inputs (reference file format plus
the sidecar), runs the reference build the case names in a scratch directory, and
stores in <case>/expected.json the SHA-256 of each of the 17 output files and of
stdout (without the wall-clock line).  Cases with ``dump`` also keep the final state
at full precision in <case>/dump.npz, written by the hooked build; for a "plain" case
the hooked build is run as well and its output files must hash equal to the plain
build's (the hooks are inert when unset).  Only data is stored: inputs, digests,
numbers.
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from cases import CASES, OUTPUT_FILES, REFERENCE_DIR, hook_env, read_dump, stdout_lines  # noqa: E402
from swehr_amd.inputs import write_case  # noqa: E402

PLAIN = os.path.join(ROOT, "oracle", "_ref", "swehr_ref")
HOOKED = os.path.join(ROOT, "oracle", "_ref", "swehr_ref_hooked")


def sha(path):
    if not os.path.exists(path):
        return None
    h = hashlib.sha256()
    with open(path, "rb") as fh:
        for blk in iter(lambda: fh.read(1 << 20), b""):
            h.update(blk)
    return h.hexdigest()


def run_ref(binary, inputs_dir, env_extra):
    d = tempfile.mkdtemp(prefix="swehr_gold_")
    for fn in os.listdir(inputs_dir):
        shutil.copy(os.path.join(inputs_dir, fn), d)
    env = dict(os.environ)
    env.update({k: (v.replace(inputs_dir, d) if isinstance(v, str) else v) for k, v in env_extra.items()})
    p = subprocess.run([binary], cwd=d, env=env, capture_output=True, text=True, timeout=3600)
    if p.returncode != 0:
        raise RuntimeError("%s failed in %s: %s" % (binary, d, p.stderr[-2000:]))
    out = {fn: sha(os.path.join(d, fn)) for fn in OUTPUT_FILES}
    lines = stdout_lines(p.stdout)
    return d, out, lines


def main(names):
    if not os.path.isdir(REFERENCE_DIR) or not os.path.exists(PLAIN) or not os.path.exists(HOOKED):
        raise SystemExit("needs %s and the reference builds (make -C oracle ref)" % REFERENCE_DIR)
    for gc in CASES:
        if names and gc.name not in names:
            continue
        case = gc.make()
        cdir = os.path.join(HERE, gc.name)
        inputs = os.path.join(cdir, "inputs")
        if gc.name != "shipped":
            if os.path.isdir(inputs):
                shutil.rmtree(inputs)
            write_case(case, inputs)
        os.makedirs(cdir, exist_ok=True)
        print("case %s (%s build, %d x %d, S=%d)" % (gc.name, gc.build, case.nx, case.ny, case.nclass), flush=True)
        if gc.build == "plain":
            if os.path.exists(os.path.join(inputs, "swehr_ext.txt")):
                raise SystemExit("plain case %s needs a sidecar" % gc.name)
            d, hashes, lines = run_ref(PLAIN, inputs, {"REFH_MAXSTEPS": str(gc.maxsteps)} if gc.maxsteps else {})
        else:
            d, hashes, lines = run_ref(HOOKED, inputs, hook_env(case, inputs, gc.maxsteps))
        shutil.rmtree(d)
        rec = dict(name=gc.name, build=gc.build, maxsteps=gc.maxsteps, note=gc.note, files=hashes,
                   stdout_sha256=hashlib.sha256("\n".join(lines).encode()).hexdigest(),
                   stdout_lines=len(lines), stdout_head=lines[:20], stdout_tail=lines[-3:])
        if gc.dump:
            dump_dir = tempfile.mkdtemp(prefix="swehr_dump_")
            dpath = os.path.join(dump_dir, "dump.bin")
            d2, hashes2, lines2 = run_ref(HOOKED, inputs, hook_env(case, inputs, gc.maxsteps, dpath))
            shutil.rmtree(d2)
            if hashes2 != hashes or lines2 != lines:
                raise SystemExit("hooked build differs from the %s build on %s" % (gc.build, gc.name))
            rec["hooked_equal"] = True
            dump = read_dump(dpath)
            shutil.rmtree(dump_dir)
            rec["steps"] = dump["steps"]
            rec["t"] = dump["t"]
            rec["stagecnt"] = dump["stagecnt"]
            arrays = {k: v for k, v in dump.items() if isinstance(v, np.ndarray)}
            np.savez_compressed(os.path.join(cdir, "dump.npz"), **arrays)
        with open(os.path.join(cdir, "expected.json"), "w") as fh:
            json.dump(rec, fh, indent=1, sort_keys=True)
            fh.write("\n")
        print("  steps=%s stdout=%d lines" % (rec.get("steps"), len(lines)), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
