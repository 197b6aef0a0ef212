"""The `abm` drop-in: command-line surface of the reference (CPU-only checks;
the run itself is covered on the GPU by test_gpu_abm.py)."""
import os
import subprocess

import pytest

import paragraph_amd as pg


def write_inputs(d):
    open(os.path.join(d, "nodes.csv"), "w").write("#node,year\n10,2000\n20,2001\n30,2001\n40,2002\n")
    open(os.path.join(d, "edges.csv"), "w").write("20,10\n30,10\n40,20\n40,99\n")
    open(os.path.join(d, "bag.csv"), "w").write("0,3\n")
    open(os.path.join(d, "rec.csv"), "w").write("0,1\n")


def base_args(d):
    j = lambda f: os.path.join(d, f)
    return ["--edgelist", j("edges.csv"), "--nodelist", j("nodes.csv"), "--out-degree-bag",
            j("bag.csv"), "--recency-probabilities", j("rec.csv"), "--growth-rate", "0.5",
            "--num-cycles", "2", "--same-year-proportion", "0.12", "--output-file", j("out.csv"),
            "--auxiliary-information-file", j("aux.csv"), "--log-file", j("log.txt")]


def run(args):
    return subprocess.run([pg.ABM_PATH] + args, capture_output=True, text=True, timeout=120)


def test_driver_command_line_parses(tmp_path):
    """Flags of drivers/parallel_gpu_10y_static.sh:41-64 and drivers/run10y3p.sh:3."""
    d = str(tmp_path)
    write_inputs(d)
    extra = ["--alpha", "0.5", "--preferential-weight", "0.33", "--recency-weight", "0.33",
             "--fitness-weight", "0.33", "--fully-random-citations", "0.05", "--use-warp-bfs",
             "true", "--use-batching", "1", "--max-batch-size", "20000",
             "--use-multistage-kernel", "yes", "--num-processors", "4", "--log-level", "1"]
    r = run(base_args(d) + extra)
    if os.path.exists("/dev/kfd"):
        assert r.returncode == 0, r.stderr
    else:
        # parsing and input reading succeed; only the device is missing
        assert r.returncode == 1 and "device" in r.stderr, r.stderr
    assert "not used by this growth model" in r.stderr


@pytest.mark.parametrize("bad", [["--bogus", "1"], ["--num-cycles"]])
def test_bad_arguments_fail(tmp_path, bad):
    d = str(tmp_path)
    write_inputs(d)
    r = run(base_args(d) + bad)
    assert r.returncode == 1


def test_missing_required_flag_fails(tmp_path):
    d = str(tmp_path)
    write_inputs(d)
    args = base_args(d)
    i = args.index("--growth-rate")
    del args[i:i + 2]
    r = run(args)
    assert r.returncode == 1 and "--growth-rate" in r.stderr


def test_variable_out_degree_bag_is_rejected(tmp_path):
    d = str(tmp_path)
    write_inputs(d)
    open(os.path.join(d, "bag.csv"), "w").write("0,3\n1,5\n")
    r = run(base_args(d))
    assert r.returncode == 1 and "citations-per-paper" in r.stderr


def test_reference_model_accepts_variable_bag(tmp_path):
    """--model reference takes the reference's multi-value bag and its
    defaults (src/main.cu:61-85); only the device is missing on the CPU."""
    d = str(tmp_path)
    write_inputs(d)
    open(os.path.join(d, "bag.csv"), "w").write("0,3\n1,5\n2,300\n")
    open(os.path.join(d, "rec.csv"), "w").write("#d,p\n1,0.5\n2,0.25\n")
    r = run(base_args(d) + ["--model", "reference", "--use-warp-bfs", "false"])
    assert "clamped to 250" in r.stderr and "has no effect here" in r.stderr
    if os.path.exists("/dev/kfd"):
        assert r.returncode == 0, r.stderr
    else:
        assert r.returncode == 1 and "device" in r.stderr, r.stderr


def test_reference_model_rejects_bad_tables(tmp_path):
    d = str(tmp_path)
    write_inputs(d)
    open(os.path.join(d, "rec.csv"), "w").write("2000,0.5\n")
    r = run(base_args(d) + ["--model", "reference"])
    assert r.returncode == 1 and "recency table" in r.stderr
    open(os.path.join(d, "rec.csv"), "w").write("1,0.5\n")
    r = run(base_args(d) + ["--model", "reference", "--alpha", "1.5"])
    assert r.returncode == 1 and "--alpha" in r.stderr
