"""Host-side tile arithmetic of the batched launches (mrg8_device.hpp:
segment_head / segment_tiles, shared by the host's tile map and the kernels'
tile bounds), compiled with g++ and checked exhaustively on small cases:
the tiles of a segment cover its outputs exactly once, in order, every tile
after the first starts on a boundary of the tile alignment (the largest power
of two dividing tile * element size), and no tile is empty or longer than
the head (shorter than one tile) plus one tile."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <cstdio>
#include <cstdint>
#include "mrg8_device.hpp"
using namespace mrg8;
int main() {
  long checked = 0;
  for (int fmt = 0; fmt < 3; ++fmt) {
    const uint64_t esz = fmt == 2 ? 4 : 8;
    for (uint64_t tile : {256ull, 512ull, 768ull, 4096ull}) {
      const uint64_t A = tile_align_bytes(fmt, tile);
      if (A & (A - 1) || (tile * esz) % A || ((tile * esz / A) & 1) == 0) return printf("alignment %llu\n", (unsigned long long)A), 1;
      const uint64_t misalign[] = {0, esz, 2 * esz, 120, 128, A / 2 + esz, A - esz};
      for (uint64_t mis : misalign) {
        const void* out = reinterpret_cast<const void*>(uintptr_t(1) << 30 | mis);
        for (uint64_t count = 0; count < 3 * tile + 70; count += (count < 300 ? 1 : 37)) {
          const uint64_t head = segment_head(out, fmt, count, tile);
          if (head >= tile && head != count) return printf("head too long\n"), 1;
          const uint64_t nt = segment_tiles(out, fmt, count, tile);
          if (count == 0) { if (nt != 0) return printf("empty segment has tiles\n"), 1; continue; }
          uint64_t next = 0;
          for (uint64_t kt = 0; kt < nt; ++kt) {
            uint64_t first = 0, end = 0;
            tile_bounds(head, count, tile, kt, &first, &end);  // the kernels' tile bounds (lookup_tile)
            if (first != next || end <= first) return printf("gap/empty fmt %d mis %llu count %llu kt %llu\n", fmt, (unsigned long long)mis, (unsigned long long)count, (unsigned long long)kt), 1;
            if (end - first > head + tile) return printf("too long\n"), 1;
            if (kt > 0 && ((uintptr_t(out) + first * esz) % A) != 0) return printf("misaligned tile\n"), 1;
            next = end;
            ++checked;
          }
          if (next != count) return printf("coverage fmt %d mis %llu count %llu\n", fmt, (unsigned long long)mis, (unsigned long long)count), 1;
        }
      }
    }
  }
  printf("ok %ld\n", checked);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"),
                    reason="needs g++ and the HIP headers")
def test_segment_tiles_cover_exactly(tmp_path):
    src = tmp_path / "tiles.cpp"
    src.write_text(PROG)
    exe = tmp_path / "tiles"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I" + os.path.join(ROOT, "mrg8_amd", "csrc"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
