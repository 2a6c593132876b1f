"""LDS layouts of the fixed FFT plans (StaticPlan in sassena_amd/csrc/sx_fft_device.h).

Each pass of a fixed plan stores its outputs through StaticPlan::swz and the
next pass reads them back through the same function, with addresses formed as
a per-lane base plus a constant.  The tables here come from the shipped header
itself (compiled host-only), and the checks are the identities the kernels
rely on: swz is a permutation of the half-length, the writers' and readers'
base-plus-constant addresses equal swz of the full index, and a swizzled store
of one wavefront covers every float2 position of a 32-bank row.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLANS = ["Plan256", "Plan512", "Plan1024", "Plan1536", "Plan2048", "Plan4096", "Plan5120", "Plan2560",
         "Plan3072", "Plan6144", "Plan8192", "Plan10240"]

DUMP = r"""
#include <cstdio>
#include "sx_fft_device.h"
template <typename PT>
void dump() {
    std::printf("plan %u %u %d", PT::Lh, PT::T, PT::npass);
    for (int p = 0; p < PT::npass; ++p) std::printf(" %d", PT::radix(p));
    std::printf("\n");
    for (int p = 0; p < PT::npass; ++p) {
        std::printf("pass %d %d", p, PT::swz_n(p));
        for (unsigned o = 0; o < PT::Lh; ++o) std::printf(" %u", PT::swz(p, o));
        std::printf("\n");
    }
}
int main() {
@CALLS@
    return 0;
}
"""


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


@pytest.fixture(scope="module")
def layouts(tmp_path_factory):
    cc = hipcc()
    if cc is None:
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("swz")
    src = d / "dump.hip"
    src.write_text(DUMP.replace("@CALLS@", "\n".join(f"    dump<sx::{p}>();" for p in PLANS)))
    exe = d / "dump"
    subprocess.run([cc, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-w",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "sassena_amd", "csrc"),
                    "-o", str(exe), str(src)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    plans, cur = [], None
    for line in out:
        f = line.split()
        if not f:
            continue
        if f[0] == "plan":
            Lh, T, npass = int(f[1]), int(f[2]), int(f[3])
            cur = {"Lh": Lh, "T": T, "radix": [int(x) for x in f[4:4 + npass]], "swz": [], "n": []}
            plans.append(cur)
        else:
            cur["n"].append(int(f[2]))
            cur["swz"].append([int(x) for x in f[3:]])
    assert len(plans) == len(PLANS)
    return plans


def ns_of(radix):
    ns = [1]
    for r in radix[:-1]:
        ns.append(ns[-1] * r)
    return ns


def test_swizzle_is_a_permutation(layouts):
    for P in layouts:
        for sw in P["swz"]:
            assert sorted(sw) == list(range(P["Lh"]))
        # the last pass's buffer is never read back: no layout of its own
        assert P["n"][-1] == 0


def test_swizzled_passes(layouts):
    # the passes with Ns < 16 and power-of-two radices get a layout of their own
    got = {P["Lh"]: [p for p, n in enumerate(P["n"]) if n > 0] for P in layouts}
    assert got[5120] == [0, 1] and got[10240] == [0, 1] and got[2560] == [0, 1]
    assert got[4096] == [0, 1] and got[8192] == [0, 1] and got[256] == [1]
    assert got[512] == [] and got[1536] == []


def test_writer_and_reader_addresses(layouts):
    for P in layouts:
        Lh, T, radix = P["Lh"], P["T"], P["radix"]
        ns, npass, E = ns_of(radix), len(radix), Lh // T
        sw = P["swz"]
        for p in range(npass):
            R = radix[p]
            if p >= 1:
                # readers of pass p - 1's buffer: lane base swz(ut) plus i T + q Lh / R (and q T)
                prev = sw[p - 1]
                offs = {i * T + q * (Lh // R) for i in range(E // R) for q in range(R)}
                offs |= {q * T for q in range(E)}
                for ut in range(T):
                    for C in offs:
                        if ut + C < Lh:
                            assert prev[ut] + C == prev[ut + C], (Lh, p, ut, C)
            if p == 0:
                for ut in range(T):
                    for i in range(E // R):
                        for q in range(R):
                            a = sw[0][ut * R + (q & ~1)] + (q & 1) + i * T * R
                            assert a == sw[0][ut * R + q + i * T * R], (Lh, ut, i, q)
            elif P["n"][p] > 0:
                Ns = ns[p]
                for ut in range(T):
                    kk = ut & (Ns - 1)
                    ob = (ut - kk) * R + kk
                    for i in range(E // R):
                        for q in range(R):
                            assert sw[p][ob + q * Ns] + i * T * R == sw[p][ob + q * Ns + i * T * R]


def test_swizzled_stores_cover_every_bank_position(layouts):
    for P in layouts:
        T, radix = P["T"], P["radix"]
        ns = ns_of(radix)
        for p, n in enumerate(P["n"]):
            if n == 0:
                continue
            R, Ns = radix[p], ns[p]
            for w in range(T // 64):
                lanes = range(64 * w, 64 * w + 64)
                if p == 0:
                    # 16-byte stores: pairs (q, q + 1) at even positions
                    for qe in range(0, R, 2):
                        pos = {P["swz"][0][ut * R + qe] % 16 for ut in lanes}
                        assert pos == set(range(0, 16, 2)), (P["Lh"], qe)
                else:
                    for q in range(R):
                        pos = {P["swz"][p][(ut - (ut & (Ns - 1))) * R + (ut & (Ns - 1)) + q * Ns] % 16 for ut in lanes}
                        assert pos == set(range(16)), (P["Lh"], p, q)
