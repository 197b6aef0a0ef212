"""The backward-CSR merge's device source run on the host (CPU).

oracle/xcheck/merge_host.cpp includes paragraph_amd/csrc/k_merge.inc as it is
compiled for the GPU and runs every block's 256 threads as host threads with
real barriers, wavefront exchanges and LDS atomics, at 1024, 2048 and 4096
output positions per block, with and without the partition pass, on graphs
with hubs and long runs of uncited papers.  The rows must equal the oracle's
or_csr_append.  (4096 positions gave rows that differ on the GPU in an
earlier build; this shows the source itself merges correctly there.)
"""
import os
import subprocess

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def test_merge_source_equals_oracle_at_every_block_size():
    subprocess.run(["make", "-s", "-C", ORACLE, "merge_host"], check=True)
    r = subprocess.run([os.path.join(ORACLE, "merge_host"), "small"], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("equal to the oracle") == 18
