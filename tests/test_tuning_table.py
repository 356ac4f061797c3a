"""The tuning table the library reads at run time (csrc/tuning.hip): parsing, nearest-entry
lookup in (log nodes, log particles), defaults, and the committed table's entries.  Host code
only: runs without a GPU."""
import os

import pic_amd as P

PKG = os.path.dirname(os.path.abspath(P.__file__))


def test_nearest_entry_and_defaults(tmp_path):
    t = tmp_path / "table.txt"
    t.write_text("# comment line\n"
                 "push 1000 1000000 chunk=1024 stripe=4  # small\n"
                 "push 1000000 1000000000 chunk=65536 stripe=16\n"
                 "mg 1000000 0 sweeps=2\n")
    P.tuning_load(str(t))
    try:
        assert P.tuning_lookup("push", 2000, 3e6, "chunk", -1) == 1024
        assert P.tuning_lookup("push", 5e5, 5e8, "stripe", -1) == 16
        assert P.tuning_lookup("mg", 1e3, 0, "sweeps", -1) == 2  # the only mg entry
        assert P.tuning_lookup("mg", 1e6, 0, "gal_from", 7) == 7  # key absent: default
        assert P.tuning_lookup("ensemble", 1e6, 1e8, "chunk", 4096) == 4096  # kind absent
    finally:
        P.tuning_load(os.path.join(PKG, "tuning", "table.txt"))


def test_committed_table_has_every_kind():
    path = os.path.join(PKG, "tuning", "table.txt")
    kinds = {line.split()[0] for line in open(path) if line.strip() and not line.startswith("#")}
    assert {"push", "mg", "ensemble"} <= kinds
    P.tuning_load(path)
    # the sphere run's sizes (257x257x513 nodes, 1e9 particles) find the push entry
    chunk = P.tuning_lookup("push", 257 * 257 * 513, 1.02e9, "chunk", -1)
    assert chunk >= 256 and chunk != -1
