"""Writes chain_asm.h: sequential FP64 sums of n doubles in LDS (n = 512, 64, 16) as one inline
assembly block each.

A sequential sum is bound by its dependent FP64 adds (5.7 clocks each on MI355X, measured in a
register-only chain) and by the instructions that bring the values in: a single wave pays
about 11 clocks to issue one LDS read and 5 clocks per s_waitcnt.  Measured per add, one wave,
eight chains in eight lanes: the compiler's own code (two or three 128-bit reads ahead) ~16;
64-bit reads with a wait per value 16.3; 128-bit reads with a wait per read 10.4; 128-bit
reads waited for and issued in groups of GROUP, INFLIGHT outstanding, 8.1.

Each block keeps INFLIGHT 128-bit reads outstanding in a ring of INFLIGHT + GROUP register
quads at fixed registers v[BASE : BASE + 4 * (INFLIGHT + GROUP) - 1] (declared clobbered, so
the compiler keeps nothing there across the block).  A step waits for the oldest GROUP reads,
issues the next GROUP reads, then adds the 2 * GROUP values in order; every read is waited for
inside the block (the last step waits for all), so no load is in flight when the block ends.
BASE is a template argument: 16 for kernels that chain at the end of a short body, 200 in
k_bottom, whose Krylov state occupies the low registers.

    python sayakaamr_amd/csrc/dev/gen_chain_asm.py   # rewrites chain_asm.h beside this script
"""
import os

INFLIGHT = 8
GROUP = 4
RING = INFLIGHT + GROUP
SIZES = (512, 64, 16)
BASES = (16, 200)


def block(n, base):
    nr = n // 2                       # 128-bit reads
    quad = lambda j: f"v[{base + 4 * (j % RING)}:{base + 4 * (j % RING) + 3}]"
    lo = lambda j: f"v[{base + 4 * (j % RING)}:{base + 4 * (j % RING) + 1}]"
    hi = lambda j: f"v[{base + 4 * (j % RING) + 2}:{base + 4 * (j % RING) + 3}]"
    lines = ["s_waitcnt lgkmcnt(0)"]
    issued = 0
    while issued < min(INFLIGHT, nr):
        lines.append(f"ds_read_b128 {quad(issued)}, %[a] offset:{16 * issued}")
        issued += 1
    i = 0
    while i < nr:
        g = min(GROUP, nr - i)
        lines.append(f"s_waitcnt lgkmcnt({issued - (i + g)})")   # reads i .. i+g-1 have landed
        for _ in range(g):
            if issued < nr:
                lines.append(f"ds_read_b128 {quad(issued)}, %[a] offset:{16 * issued}")
                issued += 1
        for j in range(i, i + g):
            lines.append(f"v_add_f64 %[s], %[s], {lo(j)}")
            lines.append(f"v_add_f64 %[s], %[s], {hi(j)}")
        i += g
    return lines


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    out = [
        "// chain_asm.h -- generated by gen_chain_asm.py (same folder); do not edit by hand.",
        "//",
        "// lds_chain_asm<N, BASE>(addr, s): s + v[0] + v[1] + ... + v[N-1], added one at a time in",
        "// index order (the same sequence of roundings as a C loop), the N doubles at LDS byte address",
        f"// addr (16-byte aligned).  {INFLIGHT} 128-bit LDS reads stay in flight in a ring of register",
        f"// quads at v[BASE .. BASE + {4 * RING - 1}], waited for and reissued {GROUP} at a time; every read",
        "// is waited for inside the block.",
        "#pragma once",
        "",
        "namespace sydev {",
        "",
        "template <int N, int BASE>",
        "__device__ inline double lds_chain_asm(unsigned addr, double s);",
        "",
    ]
    for base in BASES:
        clob = ", ".join(f'"v{r}"' for r in range(base, base + 4 * RING))
        for n in SIZES:
            out.append("template <>")
            out.append(f"__device__ inline double lds_chain_asm<{n}, {base}>(unsigned addr, double s) {{")
            out.append("    asm volatile(")
            for ln in block(n, base):
                out.append(f'        "{ln}\\n"')
            out.append('        : [s] "+v"(s)')
            out.append('        : [a] "v"(addr)')
            out.append(f'        : "memory", {clob});')
            out.append("    return s;")
            out.append("}")
            out.append("")
    out.append("}  // namespace sydev")
    with open(os.path.join(here, "chain_asm.h"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
