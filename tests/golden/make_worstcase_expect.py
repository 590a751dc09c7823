"""Write worstcase-110-terms-4-m.k3.json: the expected full version of every
relation of the reference's tests/worstcase-110-terms-4-m.cpp (compiled
unchanged against the facade and linked with the oracle's implementation of
the C ABI, oracle/libpra_abi_oracle.so) at 3 buckets, as row counts and
sha256 digests of the tsv_to_bin dump.  Inputs: the reference's
data/worstcase-110-terms-4-m files, copied next to this script.

Run from the repository root after building the oracle (make -C oracle all):
    python tests/golden/make_worstcase_expect.py
"""
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = "/root/reference/tests/worstcase-110-terms-4-m.cpp"
NB = 3


def main():
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.dirname(HERE))
    import __graft_entry__ as ge
    from programs import dump_digest, run_facade_program, total_facts
    with tempfile.TemporaryDirectory() as d:
        exe = ge.build_reference_program(SRC, os.path.join(d, "wc_oracle"), libdir=ge.ORACLE, lib="pra_abi_oracle")
        r = run_facade_program(exe, "worstcase-110-terms-4-m", os.path.join(d, "run"), NB, os.path.join(d, "dump"))
        assert r.returncode == 0, r.stderr
        dig = dump_digest(os.path.join(d, "dump"))
    out = {"program": "tests/worstcase-110-terms-4-m.cpp", "buckets": NB, "total_facts": total_facts(r.stdout),
           "relations": dig}
    with open(os.path.join(HERE, "worstcase-110-terms-4-m.k3.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(out["total_facts"], len(dig))


if __name__ == "__main__":
    main()
