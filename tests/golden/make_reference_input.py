"""Writes reference_read_input.json: the grid, physics and run settings that
the library's parser (mlg_read_input, mlg_read_run_settings) reads from the
reference's own read.input, so that GPU tests can set up the reference's
shipped case without the reference tree.  Run from the repository root
where /root/reference exists: python tests/golden/make_reference_input.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mlegs_amd import hip as H  # noqa: E402

SRC = "/root/reference/read.input"


def main():
    params, t_limit = H.read_input(SRC)
    st = H.read_run_settings(SRC)
    keep = {k: st[k] for k in ("rmv_sw", "rmv_int", "adj_sw", "adj_int", "it1", "it2") if k in st}
    out = {"source": "read.input of the reference (parsed by mlg_read_input / mlg_read_run_settings)",
           "params": params, "t_limit": t_limit, "run_settings": keep}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_read_input.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
