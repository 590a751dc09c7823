"""Write tc_g13236_k2_checkpoints.json: sha256 and size of every file of the
checkpoints that the reference's tests/transitive_closure.cpp (compiled
unchanged against the facade, linked with the oracle's implementation of the
C ABI, oracle/libpra_abi_oracle.so) writes for data/g13236 at 2 buckets with
a checkpoint every loop, plus the dumped relations after a restart from
checkpoint-1-5.

Run from the repository root after building the oracle (make -C oracle all):
    python tests/golden/make_checkpoint_expect.py
"""
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = "/root/reference/tests/transitive_closure.cpp"


def main():
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.dirname(HERE))
    import __graft_entry__ as ge
    from programs import checkpoint_digest, dump_digest, run_tc_checkpointed
    with tempfile.TemporaryDirectory() as d:
        exe = ge.build_reference_program(SRC, os.path.join(d, "tc_oracle"), libdir=ge.ORACLE, lib="pra_abi_oracle")
        out, restart_dump = run_tc_checkpointed(exe, d)
        exp = {"program": "tests/transitive_closure.cpp", "edges": "g13236_edge_2_1_2", "buckets": 2,
               "cp_iteration": 1, "checkpoints": checkpoint_digest(out),
               "restart_from": "checkpoint-1-5", "restart_dump": dump_digest(restart_dump)}
    with open(os.path.join(HERE, "tc_g13236_k2_checkpoints.json"), "w") as f:
        json.dump(exp, f, indent=1, sort_keys=True)
    print(len(exp["checkpoints"]), "checkpoints")


if __name__ == "__main__":
    main()
