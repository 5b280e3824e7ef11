"""The host C code that runs without a GPU — the `.tns` reader (parallel
line split and parse), empty-index removal and the constraint callbacks —
built with AddressSanitizer and UndefinedBehaviorSanitizer (host gcc, no
device code) and driven over the fixtures, malformed and empty files and
edge shapes (0x0, 1xR, Ix1). Any sanitizer report fails the test."""
import os
import shutil
import subprocess

import pytest

from conftest import DATASETS, ROOT, tensor_path

HOST = os.path.join(ROOT, "splatt-stream_amd", "host")
INCS = ["-I" + HOST, "-I" + os.path.join(ROOT, "splatt-stream_amd", "csrc"),
        "-I" + os.path.join(ROOT, "include")]

DRIVER = r"""
#include "host.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void * splatt_malloc(size_t const bytes) { return malloc(bytes); }
void splatt_free(void * p) { free(p); }

static int read_one(char const * path)
{
  sptensor_t * tt = splatt_tt_read(path);
  if(tt == NULL) { printf("%s NULL\n", path); return 0; }
  double s = 0.;
  for(idx_t n=0; n < tt->nnz; ++n) {
    s += tt->vals[n];
    for(idx_t m=0; m < tt->nmodes; ++m) s += (double) tt->ind[m][n];
  }
  idx_t const removed = splatt_tt_remove_empty(tt);
  printf("%s nnz %lu modes %lu removed %lu sum %.17g\n", path,
      (unsigned long) tt->nnz, (unsigned long) tt->nmodes, (unsigned long) removed, s);
  splatt_tt_free(tt);
  return 0;
}

static void prox_all(void)
{
  idx_t const shapes[][2] = {{0, 0}, {1, 1}, {1, 5}, {5, 1}, {2, 3}, {37, 16}, {300, 64}};
  double lam = 0.3;
  srand(3);
  for(unsigned s=0; s < sizeof(shapes) / sizeof(shapes[0]); ++s) {
    idx_t const I = shapes[s][0], R = shapes[s][1];
    for(int k=0; k < 7; ++k) {
      double * P = (double *) malloc((I * R ? I * R : 1) * sizeof(double));
      for(idx_t x=0; x < I * R; ++x) P[x] = rand() / (double) RAND_MAX - 0.4;
      switch(k) {
        case 0: splatt_maxcolnorm_prox(P, I, R, 0, NULL, 1.5, false); break;
        case 1: splatt_maxcolnorm_nonneg_prox(P, I, R, 0, NULL, 1.5, false); break;
        case 2: splatt_nonneg_prox(P, I, R, 0, NULL, 1.5, false); break;
        case 3: splatt_lasso_prox(P, I, R, 0, &lam, 1.5, false); break;
        case 4: splatt_rowsimp_prox(P, I, R, 0, NULL, 1.5, false); break;
        case 5: splatt_smooth_prox(P, I, R, 0, &lam, 1.5, false); break;
        default: splatt_colsimp_prox(P, I, R, 0, NULL, 1.5, false); break;
      }
      free(P);
    }
  }
  printf("prox ok\n");
}

int main(int argc, char ** argv)
{
  for(int a=1; a < argc; ++a) read_one(argv[a]);
  prox_all();
  return 0;
}
"""

EDGE_FILES = {
    "comments.tns": "# c\n1 2 3 1.5\n\n2 1 1 2e0\n#x\n3 3 2 -1",
    "empty.tns": "",
    "only_comments.tns": "\n\n#only\n",
    "zeroidx.tns": "0 0 0.5\n1 4 1.25\n",
    "tabs.tns": "1\t2\t3\n\t\n",
    "short_line.tns": "1 2 3 4\n1 2\n",
    "bad_value.tns": "1 2 x\n",
}


def _big_file(path, lines):
    with open(path, "w") as f:
        for i in range(lines):
            f.write("%d %d %d %.6f\n" % (1 + i % 997, 1 + (i * 7) % 1009, 1 + (i * 13) % 31,
                                         0.5 + (i % 100) / 7.0))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs the host gcc")
def test_host_reader_and_prox_sanitized(tmp_path):
    exe = str(tmp_path / "host_san")
    drv = tmp_path / "driver.c"
    drv.write_text(DRIVER)
    cmd = (["gcc", "-g", "-O1", "-std=gnu11", "-fopenmp", "-fsanitize=address,undefined",
            "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"] + INCS +
           [str(drv), os.path.join(HOST, "coord.c"), os.path.join(HOST, "opts.c"),
            "-o", exe, "-lm"])
    subprocess.run(cmd, check=True, capture_output=True, timeout=120)

    files = []
    for name, text in EDGE_FILES.items():
        p = tmp_path / name
        p.write_text(text)
        files.append(str(p))
    big = str(tmp_path / "big.tns")
    _big_file(big, 50000)
    files.append(big)
    files += [tensor_path(d) for d in DATASETS]

    env = dict(os.environ, OMP_NUM_THREADS="8",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe] + files, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    out = {l.split()[0]: l for l in r.stdout.splitlines() if l.strip()}
    assert out[files[0]].split()[1:5] == ["nnz", "3", "modes", "3"]
    assert "NULL" in out[files[1]] and "NULL" in out[files[2]]
    assert out[big].split()[1:3] == ["nnz", "50000"]
    assert "prox ok" in r.stdout
