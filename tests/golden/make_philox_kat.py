"""Write tests/golden/philox_kat.json: Philox4x32-10 known answers.

The answers come from rocRAND's header-only Philox4x32-10 engine
(/opt/rocm/include/rocrand/rocrand_philox4x32_10.h, read as source and compiled
here for the host), an implementation independent of both the oracle and the
device kernel.  Run once in the build container:

    python tests/golden/make_philox_kat.py
"""
import json
import os
import subprocess
import tempfile

CASES = [
    # (counter[4], key[2]) — edge patterns plus counters shaped like the ones the
    # propagation uses: (block, photon_lo, photon_hi, stream).
    ((0, 0, 0, 0), (0, 0)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)),
    ((1, 0, 0, 0), (0, 0)),
    ((0, 1, 0, 0), (0, 0)),
    ((0, 0, 1, 0), (0, 0)),
    ((0, 0, 0, 1), (0, 0)),
    ((0, 0, 0, 0), (1, 0)),
    ((0, 0, 0, 0), (0, 1)),
    ((0, 12345, 0, 0), (0, 0)),
    ((3, 999999999, 0, 7), (42, 0)),
    ((8, 0xDEADBEEF, 0x1, 4095), (0x9E3779B9, 0x7F4A7C15)),
    ((2, 1000000, 0, 0), (0xFFFFFFFF, 0)),
    ((5, 0x80000000, 0x3, 0xABCDEF), (0x12345678, 0x9ABCDEF0)),
]

SRC = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct P : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};
int main() {
    P p;
    unsigned int c0, c1, c2, c3, k0, k1;
    while (std::scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 o = p.run(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        std::printf("%u %u %u %u\n", o.x, o.y, o.z, o.w);
    }
    return 0;
}
"""


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "kat.cpp")
        exe = os.path.join(td, "kat")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.check_call(["hipcc", "-O1", "-o", exe, src])
        inp = "".join("%d %d %d %d %d %d\n" % (*c, *k) for c, k in CASES)
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    rows = []
    for (c, k), line in zip(CASES, out.strip().splitlines()):
        rows.append({"counter": list(c), "key": list(k), "out": [int(x) for x in line.split()]})
    with open(os.path.join(here, "philox_kat.json"), "w") as f:
        json.dump({"source": "rocRAND rocrand_philox4x32_10.h ten_rounds (ROCm 7.2)",
                   "cases": rows}, f, indent=1)
    print("wrote", len(rows), "cases")


if __name__ == "__main__":
    main()
