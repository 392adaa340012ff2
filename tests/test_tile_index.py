"""Host-side pin of the bracket kernel's index arithmetic.

The kernel takes its tile origin, halo rows / columns and periodic wraps from
vivlasov1d_amd/csrc/tile_index.h.  tests/native/tile_index_check.cpp compiles
that header for the host and enumerates, for every grid nx in [8, 130] x
nv in [8, 70] and the benchmark / reference-example sizes, every tile, every
staged halo cell, every x- and v-factor load and every stencil read: each must
be inside its buffer and be the periodic-in-x neighbour the stencil needs.
An over-read in a halo cell that feeds no stored output passes every parity
test; this one fails on it (a single periodic wrap, the fault of an earlier
version on nx < 66, reports thousands of violations)."""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "tile_index_check.cpp")


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    pytest.skip("hipcc not available")


def test_bracket_tile_indices_in_bounds(tmp_path):
    exe = str(tmp_path / "tile_index_check")
    subprocess.run([_hipcc(), "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", SRC, "-o", exe],
                   check=True, cwd=str(tmp_path))
    sizes = ["1024", "2048", "256", "512", "64", "128", "100", "200", "50", "201", "51", "101", "9", "9",
             "4096", "16", "16", "4096", "512", "1024"]
    out = subprocess.run([exe, "8", "130", "8", "70"] + sizes, capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "0 failures" in out.stdout


def test_fused_jac_rfft_indices_in_bounds(tmp_path):
    # the fused J u + rFFT kernel's row blocks, strips, halo cells, stencil
    # window and FFT-buffer stores for nx = 256, 512, 1024, 2 / 4 / 8 rows per
    # workgroup, nv in [8, 80] and the benchmark / reference sizes
    exe = str(tmp_path / "tile_index_check")
    subprocess.run([_hipcc(), "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", SRC, "-o", exe],
                   check=True, cwd=str(tmp_path))
    out = subprocess.run([exe, "jr", "8", "80", "128", "200", "201", "256", "509", "512", "1024", "2048"],
                         capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert "0 failures" in out.stdout


# (text in tile_index.h, its replacement, checker arguments that must report it)
PLANTED = {
    'a single periodic wrap': ('    if (gi >= nx) gi %= nx;\n', '', ["8", "130", "8", "70"]),
    'bracket halo column off by one': ('return wrap_x(t.i0 - 2 + c, nx);\n}', 'return wrap_x(t.i0 - 1 + c, nx);\n}',
                                       ["8", "130", "8", "70"]),
    'fused strip column off by one': ('return wrap_x(cb - 2 + c, nx); }', 'return wrap_x(cb - 1 + c, nx); }',
                                      ["jr", "8", "80", "1024", "2048"]),
    'fused right halo columns shifted': ('*col = hc < 2 ? hc : kJrSW + hc;', '*col = hc < 2 ? hc : kJrSW + hc - 1;',
                                         ["jr", "8", "80", "1024", "2048"]),
}


@pytest.mark.parametrize("name", sorted(PLANTED))
def test_planted_index_mistake_is_reported(name, tmp_path):
    # the checker built against a copy of the header with one mistake
    old, new, args = PLANTED[name]
    text = open(os.path.join(ROOT, "vivlasov1d_amd", "csrc", "tile_index.h")).read()
    assert text.count(old) == 1, name
    (tmp_path / "vivlasov1d_amd" / "csrc").mkdir(parents=True)
    (tmp_path / "tests" / "native").mkdir(parents=True)
    (tmp_path / "vivlasov1d_amd" / "csrc" / "tile_index.h").write_text(text.replace(old, new))
    src = tmp_path / "tests" / "native" / "tile_index_check.cpp"
    src.write_text(open(SRC).read())
    exe = str(tmp_path / "planted_check")
    subprocess.run([_hipcc(), "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", str(src), "-o", exe],
                   check=True, cwd=str(tmp_path))
    out = subprocess.run([exe] + args, capture_output=True, text=True)
    m = re.search(r"checked \d+ grids, (\d+) failures", out.stdout)
    assert out.returncode != 0 and m and int(m.group(1)) > 0, out.stdout[-300:]
