"""Philox4x32-10 known-answer vectors from rocRAND's header-only engine
(/opt/rocm/include/rocrand/rocrand_philox4x32_10.h), compiled for the
host with hipcc.  Writes philox_kat.json (data only)."""
import json
import os
import random
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

SRC = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
    Probe p;
    unsigned c0, c1, c2, c3, k0, k1;
    while (scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 r = p.run(make_uint4(c0, c1, c2, c3), make_uint2(k0, k1));
        printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
"""


def main():
    rnd = random.Random(2024)
    cases = [(0, 0, 0, 0, 0, 0),
             (0xFFFFFFFF,) * 6,
             (0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344, 0xA4093822, 0x299F31D0)]
    cases += [tuple(rnd.getrandbits(32) for _ in range(6)) for _ in range(61)]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "kat.cpp")
        exe = os.path.join(d, "kat")
        open(src, "w").write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-o", exe, src], check=True)
        inp = "".join("%x %x %x %x %x %x\n" % c for c in cases)
        out = subprocess.run([exe], input=inp, capture_output=True, text=True,
                             check=True).stdout.split("\n")
    vectors = []
    for c, line in zip(cases, out):
        vectors.append({"ctr": list(c[:4]), "key": list(c[4:]),
                        "out": [int(w, 16) for w in line.split()]})
    json.dump({"source": "rocrand_philox4x32_10.h ten_rounds", "vectors": vectors},
              open(os.path.join(HERE, "philox_kat.json"), "w"))


if __name__ == "__main__":
    main()
