"""CSRHost::partition_rows and build_owner_ghost_maps (main.cu:526-692)
through the C++ API of libmek.so (host code only, no GPU): a program compiled
against include/mek/csr_host.hpp prints the blocks and halo plans of a small
matrix, which are compared with numpy slicing and with the halo lists
written out from their definitions; the one-block partition of a matrix the
caller gives up takes its arrays over without a copy."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, gen, mek

SRC = os.path.join(ROOT, "tests", "cpp", "partition_check.cpp")


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    out = tmp_path_factory.mktemp("partition") / "partition_check"
    libdir = os.path.dirname(mek.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{os.path.join(ROOT, 'include')}", SRC,
                    "-o", str(out), f"-L{libdir}", "-lmek", f"-Wl,-rpath,{libdir}"], check=True)
    return str(out)


def _run(prog, path, arg):
    p = subprocess.run([prog, str(path), str(arg)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, (p.stdout, p.stderr)
    return [ln for ln in p.stdout.splitlines() if not ln.startswith("Binary CSR header")]


def _parse(lines, p):
    it = iter(lines)
    parts, maps = [], []
    for _ in range(p):
        h = next(it).split()
        rows, nnz = int(h[3]), int(h[7])
        rp = [int(x) for x in next(it).split()[1:]]
        ci = [int(x) for x in next(it).split()[1:]]
        va = [float(x) for x in next(it).split()[1:]]
        parts.append((rows, nnz, rp, ci, va))
    rest = list(it)
    cur = None
    for ln in rest:
        w = ln.split()
        if w[0] == "map":
            cur = {"recv": int(w[3]), "send": int(w[5]), "recv_from": {}, "send_to": {}}
            maps.append(cur)
        elif w[0] in ("recv_from", "send_to"):
            cur[w[0]][int(w[1])] = [int(x) for x in w[2:]]
        elif w[0] != "validated":
            cur[w[0]] = [int(x) for x in w[1:]]
    assert rest[-1] == "validated"
    return parts, maps


@pytest.mark.parametrize("p", [1, 2, 3, 4, 7])
def test_partitions_and_ghost_maps(tmp_path, prog, p):
    n = 23
    rp, ci, va = gen.random_rows(n, 1, 6, seed=p)
    path = tmp_path / "A.bincsr"
    gen.write_bincsr(path, rp, ci, va)
    parts, maps = _parse(_run(prog, path, p), p)
    # balanced row blocks: the first n % p blocks take one row more
    sizes = [n // p + (1 if i < n % p else 0) for i in range(p)]
    lo = np.concatenate([[0], np.cumsum(sizes)])
    for i, (rows, nnz, prp, pci, pva) in enumerate(parts):
        r0, r1 = lo[i], lo[i + 1]
        assert rows == sizes[i]
        np.testing.assert_array_equal(prp, rp[r0:r1 + 1] - rp[r0])
        np.testing.assert_array_equal(pci, ci[rp[r0]:rp[r1]])
        np.testing.assert_array_equal(pva, va[rp[r0]:rp[r1]])
        assert nnz == rp[r1] - rp[r0]
    owner = np.searchsorted(lo[1:], np.arange(n), side="right")
    for me in range(p):
        g = maps[me]
        mine = ci[rp[lo[me]]:rp[lo[me + 1]]]
        need = {o: sorted(set(int(c) for c in mine if owner[c] == o)) for o in range(p) if o != me}
        give = {}
        for q in range(p):
            if q == me:
                continue
            theirs = ci[rp[lo[q]]:rp[lo[q + 1]]]
            give[q] = sorted(set(int(c) for c in theirs if owner[c] == me))
        # every other block has an entry (possibly empty), as in the reference
        assert g["recv_from"] == need and g["send_to"] == give, me
        flat_in = [c for o in range(p) if o != me for c in need[o]]
        flat_out = [c for q in range(p) if q != me for c in give[q]]
        assert g["incoming"] == flat_in and g["outgoing"] == flat_out
        assert g["recv"] == len(flat_in) and g["send"] == len(flat_out)
        off = 0
        for o in range(p):
            if o != me and need[o]:
                assert (g["recv_offsets"][o], g["recv_counts"][o]) == (off, len(need[o]))
                off += len(need[o])
        assert len(g["recv_offsets"]) == p and len(g["send_counts"]) == p


def test_one_block_of_a_given_up_matrix_is_not_copied(tmp_path, prog):
    rp, ci, va = gen.laplacian_2d(20)
    path = tmp_path / "L.bincsr"
    gen.write_bincsr(path, rp, ci, va)
    assert _run(prog, path, "move") == ["moved 1 emptied 1", "copied 1"]
