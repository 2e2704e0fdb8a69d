"""include/forcepho_amd.h as a C caller sees it: a C99 program (no C++) that
includes the header, checks the struct sizes against the library's, links
libfpx.so and calls the host-only entry points.  No GPU needed."""

import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(REPO, "forcepho_amd")

PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "forcepho_amd.h"

#define CHECK(c) do { if (!(c)) { printf("failed: %s\n", #c); return 1; } } while (0)

int main(void) {
    int32_t exp_start[2] = {0, 256}, exp_N[2] = {256, 200};
    int16_t band_start[1] = {0}, band_N[1] = {2};
    FpxTile tiles[8];
    int n, cfg, ppt = 0;
    CHECK(fpx_abi_version() == FPX_ABI_VERSION);
    CHECK(fpx_struct_size(FPX_STRUCT_PATCH) == sizeof(FpxPatch));
    CHECK(fpx_struct_size(FPX_STRUCT_SOURCE) == sizeof(FpxSource));
    CHECK(fpx_struct_size(FPX_STRUCT_RESPONSE) == sizeof(FpxResponse));
    CHECK(fpx_struct_size(FPX_STRUCT_PSFGAUSS) == sizeof(FpxPSFSourceGaussian));
    CHECK(fpx_struct_size(FPX_STRUCT_TILE) == sizeof(FpxTile));
    CHECK(fpx_struct_size(FPX_STRUCT_SLOT) == sizeof(FpxSlot));
    CHECK(sizeof(FpxPatch) == 152 && sizeof(FpxSource) == 264 && sizeof(FpxSlot) == 88);
    CHECK(sizeof(FpxTile) == 16 && sizeof(FpxPSFSourceGaussian) == 28);
    CHECK(fpx_n_configs() > 0);
    for (cfg = 0; cfg < fpx_n_configs(); ++cfg) {
        CHECK(fpx_config_info(cfg, NULL, NULL, &ppt) == FPX_OK && ppt % FPX_WAVE == 0);
        CHECK(fpx_split_bytes(1, cfg) == 256 + 4 * (size_t)ppt);
    }
    /* 2 exposures of 256 and 200 pixels in 128-pixel tiles: 2 + 2 tiles */
    n = fpx_plan_tiles(exp_start, exp_N, band_start, band_N, 1, 128, NULL, 0);
    CHECK(n == 4);
    memset(tiles, 0xff, sizeof(tiles));
    CHECK(fpx_plan_tiles(exp_start, exp_N, band_start, band_N, 1, 128, tiles, 8) == 4);
    CHECK(tiles[3].exposure == 1 && tiles[3].band == 0 && tiles[3].pix_start == 384 &&
          tiles[3].pix_count == 72 && tiles[3].part == 0 && tiles[3].parts == 0);
    /* an error is a negative code with a message */
    CHECK(fpx_plan_tiles(exp_start, exp_N, band_start, band_N, 1, 100, NULL, 0) == FPX_E_ARG);
    CHECK(fpx_last_error() != NULL && strlen(fpx_last_error()) > 0);
    printf("ok\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_header_is_c99_and_links(tmp_path):
    lib = os.path.join(LIB_DIR, "libfpx.so")
    if not os.path.exists(lib):
        pytest.skip("libfpx.so is not built (__graft_entry__.build())")
    src = tmp_path / "caller.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "caller"
    cc = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic",
          "-I" + os.path.join(REPO, "include"), str(src), "-o", str(exe),
          "-L" + LIB_DIR, "-lfpx", "-Wl,-rpath," + LIB_DIR]
    r = subprocess.run(cc, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
