"""Writes pdep_kat.json: known answers of the x86 BMI2 instruction
_pdep_u64, which the reference's linearization uses (bitops.hpp:9-13).
Compiles a small C program with gcc -mbmi2 and records (x, mask, pdep)
for seeded random and structured masks, including the contiguous masks
(2^w - 1) << s that the key layout produces.  Run from the repo root:
    python tests/golden/make_pdep_kat.py
"""
import json
import os
import subprocess
import tempfile

SRC = r"""
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
static uint64_t sm(uint64_t *s) { uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31); }
int main(void) {
  uint64_t s = 12345; int first = 1; printf("[");
  for (int i = 0; i < 200; ++i) {
    uint64_t x = sm(&s), m = sm(&s);
    if (i % 4 == 1) { int w = (int)(sm(&s) % 33), sh = (int)(sm(&s) % (65 - w));
      m = (w == 64 ? ~0ull : ((1ull << w) - 1)) << sh; if (w == 0) m = 0; }
    if (i % 4 == 2) m &= sm(&s);
    if (i % 4 == 3) x &= 0xFFFFull;
    printf("%s[\"%llu\",\"%llu\",\"%llu\"]", first ? "" : ",", (unsigned long long)x,
           (unsigned long long)m, (unsigned long long)_pdep_u64(x, m));
    first = 0;
  }
  printf("]\n"); return 0; }
"""

def main():
    here = os.path.dirname(os.path.abspath(__file__))
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "kat.c")
        exe = os.path.join(d, "kat")
        open(c, "w").write(SRC)
        subprocess.check_call(["gcc", "-O1", "-mbmi2", "-o", exe, c])
        rows = json.loads(subprocess.check_output([exe]).decode())
    with open(os.path.join(here, "pdep_kat.json"), "w") as f:
        json.dump({"source": "x86 BMI2 _pdep_u64 (gcc -mbmi2)", "cases": rows}, f, indent=0)

if __name__ == "__main__":
    main()
