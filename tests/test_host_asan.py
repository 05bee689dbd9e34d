"""AddressSanitizer run of the command-line driver's host-memory path (-hostRead) on the CPU.

The host code (parbutterfly_amd/csrc/host.cpp: reader, validator, graph storage; csrc/cli.cpp:
flag parsing, output buffers, the per-round read / sparsify / free loop) is built with
-fsanitize=address together with tests/asan_cli_stub.cpp, which stands in for the device entry
points: it validates and reads exactly the host bytes the library's host path reads, writes
exactly the bytes it writes, and takes its results from the CPU oracle.  Every run must be
free of sanitizer reports and print the reference's totals.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CASES = ["uniform_1000x1000_m10000_s1", "rmat_8x8_m3000_s3", "k_4_5", "chunglu_2000x500_m12000_g21_s2"]


@pytest.fixture(scope="module")
def cli(tmp_path_factory):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no host compiler")
    d = tmp_path_factory.mktemp("asan")
    flags = ["-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer"]
    obj = str(d / "oracle.o")
    subprocess.run(["gcc", *flags, "-std=c99", "-c", os.path.join(ROOT, "oracle", "oracle.c"), "-o", obj], check=True)
    exe = str(d / "cli_asan")
    csrc = os.path.join(ROOT, "parbutterfly_amd", "csrc")
    subprocess.run(["g++", *flags, "-std=c++17", "-fopenmp", "-I" + os.path.join(ROOT, "include"), "-I" + csrc,
                    os.path.join(csrc, "host.cpp"), os.path.join(csrc, "cli.cpp"),
                    os.path.join(ROOT, "tests", "asan_cli_stub.cpp"), obj, "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("name", CASES)
def test_cli_host_read_under_asan(cli, name):
    adj = os.path.join(GOLD, name + ".adj")
    total = int(np.load(os.path.join(GOLD, name + ".npz"))["total_SERIAL_ADEG"][0])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", OMP_NUM_THREADS="4")
    runs = [(["-per", "VERT", "-peelType", "GPU", "-r", "1"], r"number of butterflies: (\d+)", 2),
            (["-per", "EDGE", "-peelType", "GPU", "-r", "1"], r"number of edge butterflies: (\d+)", 2),
            (["-per", "TOTAL", "-r", "2"], r"number of butterflies: (\d+)", 3),
            (["-m", "20000", "-a", "1000", "-per", "VERT", "-r", "0"], r"number of butterflies: (\d+)", 1),
            (["-sparseType", "EDGE", "-d", "3", "-r", "1"], None, 2),
            (["-sparseType", "COLOR", "-d", "2", "-r", "1"], None, 2)]
    for args, pat, rounds in runs:
        r = subprocess.run([cli, "-hostRead", *args, adj], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                           env=env, timeout=300)
        assert r.returncode == 0 and "Sanitizer" not in r.stderr, (args, r.stderr[-3000:])
        if pat:
            assert [int(x) for x in re.findall(pat, r.stdout)] == [total] * rounds, (args, r.stdout)
        else:
            assert len(re.findall(r"number of butterflies: (\d+)", r.stdout)) == rounds, (args, r.stdout)


def test_cli_timing_labels(cli):
    # every count type runs the GPU kernels; the timing line says so and names the requested type
    adj = os.path.join(GOLD, "k_4_5.adj")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", OMP_NUM_THREADS="4")
    for args, want in ((["-countType", "SERIAL", "-per", "TOTAL"], "GPU (requested Serial) : "),
                       (["-countType", "BATCHS", "-per", "VERT", "-peelType", "BATCHS"], "GPU Peel (requested Par Peel) : "),
                       (["-countType", "GPU", "-per", "EDGE"], "GPU : ")):
        r = subprocess.run([cli, "-hostRead", "-r", "0", *args, adj], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, env=env, timeout=120)
        assert r.returncode == 0 and want in r.stdout, (args, r.stdout)
        assert not re.search(r"^(Par|Serial|Par Peel) : ", r.stdout, re.M), r.stdout


def test_cli_malformed_input_under_asan(cli, tmp_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", OMP_NUM_THREADS="4")
    bad = {"empty": "", "header only": "AdjacencyHypergraph\n", "truncated": "AdjacencyHypergraph\n2\n3\n2\n3\n0\n1\n0\n",
           "range": "AdjacencyHypergraph\n1\n1\n1\n1\n0\n5\n0\n0\n", "huge": "AdjacencyHypergraph\n99999999999\n1\n1\n1\n"}
    for tag, text in bad.items():
        p = tmp_path / (tag.replace(" ", "_") + ".adj")
        p.write_text(text)
        r = subprocess.run([cli, "-hostRead", "-r", "0", str(p)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, env=env, timeout=120)
        assert r.returncode != 0 and "Sanitizer" not in r.stderr, (tag, r.stderr[-3000:])
