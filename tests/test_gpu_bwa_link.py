"""BSBolt's aligner with its two DP kernels on the GPU, end to end.

oracle/Makefile builds the reference aligner (bsbolt/External/BWA) from its own
sources twice: as is (_ref/bwa) and linked to libbsx.so's scalar drop-ins
(_ref/bwa_bsx: every ksw_extend2 call of mem_chain2aln and every ksw_global2
call of bwa_gen_cigar2 runs bsx_ksw_extend2 / bsx_ksw_global2 on the GPU).
Both align the cfg0 reads (20,000 SE 100 bp bisulfite reads on the 2 Mb cfg0
genome) with BSBolt's default argv; their SAM output must be identical, with
every DP call served by the GPU.  The binaries are built where the reference
sources are; the test is skipped where they were not built."""
import os
import subprocess
import tempfile

import pytest

import cfg0_cases as K
from conftest import ROOT

pytestmark = pytest.mark.gpu
BWA = os.path.join(ROOT, "oracle", "_ref", "bwa")
BWA_BSX = os.path.join(ROOT, "oracle", "_ref", "bwa_bsx")


def test_aligner_on_gpu_drop_ins_writes_the_reference_sam():
    if not (os.path.exists(BWA) and os.path.exists(BWA_BSX)):
        pytest.skip("reference aligner not built (oracle/Makefile needs the reference sources)")
    wl = K.workload()
    try:
        with tempfile.TemporaryDirectory() as d:
            fa, fq = K.write_fasta_fastq(d, wl)
            print("bwa index", flush=True)
            subprocess.run([BWA, "index", "-a", "bwtsw", fa], check=True, capture_output=True, timeout=120)
            sams = {}
            for name, exe in (("cpu", BWA), ("gpu", BWA_BSX)):
                print(f"bwa mem ({name})", flush=True)
                r = subprocess.run([exe, "mem", *K.BSB_ARGV, fa, fq], capture_output=True, text=True, timeout=300)
                assert r.returncode == 0, r.stderr[-2000:]
                sams[name] = [l for l in r.stdout.splitlines() if not l.startswith("@PG")]
                if name == "gpu":
                    counts = [l for l in r.stderr.splitlines() if l.startswith("bsx:")]
        assert len(sams["cpu"]) > wl.n_reads  # header + one primary record per read (+ supplementary)
        diff = [k for k, (a, b) in enumerate(zip(sams["cpu"], sams["gpu"])) if a != b]
        assert len(sams["cpu"]) == len(sams["gpu"]) and not diff, \
            f"{len(diff)} SAM lines differ, first: {sams['cpu'][diff[0]]!r} vs {sams['gpu'][diff[0]]!r}" if diff else ""
        print(counts, flush=True)
        ext = counts[-1].split()
        assert ext[1] == "ksw_extend2" and int(ext[2].split("=")[1]) >= 20000 and ext[3] == "cpu=0", counts
        assert ext[4] == "ksw_global2" and int(ext[5].split("=")[1]) > 500 and ext[6] == "cpu=0", counts
    finally:
        wl.close()
