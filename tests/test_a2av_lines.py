"""CPU model of the batched Alltoallv's line assignment (runtime/batched.h).

The copy kernel writes a row's 16-byte destination lines from several code paths: whole
lines inside one segment, lines that straddle segments (inline dword assembly, or the
straddle threads in deferred mode), and lines at a row's ends (byte path).  This model
replays the kernels' index arithmetic -- chunk rows and line ranges (scan), the 64-segment
LDS window and its coverage, the wave cursor and the four-entry selection (copy), the
owner rule of the straddle threads -- over random layouts with unaligned receive rows,
gaps, empty segments and 1/2/4/8-byte elements, and checks that every byte of every row
is written exactly once with the value MPI_Alltoallv defines (gpu_mpi/mpi.cu:363-408),
and that nothing outside the rows is written.  It runs on the CPU; the GPU tests check
the kernels themselves against the oracle.
"""
import numpy as np
import pytest

WINDOW = 64


def plan(P, elem, counts, sdisp, send_addr):
    """a2av_plan_kernel: per row d, segment end offsets, source addresses, row bytes."""
    seg_end = np.zeros((P, P), dtype=np.int64)
    seg_src = np.zeros((P, P), dtype=np.int64)
    row_bytes = np.zeros(P, dtype=np.int64)
    for d in range(P):
        carry = 0
        for s in range(P):
            c = int(counts[s, d])
            seg_end[d, s] = (carry + c) * elem
            seg_src[d, s] = send_addr[s] + int(sdisp[s, d]) * elem
            carry += c
        row_bytes[d] = carry * elem
    return seg_end, seg_src, row_bytes


def scan(P, row_bytes, recv_addr, C):
    """a2av_scan_kernel: chunks per row, counted from the 128-byte block of the first byte."""
    off = np.zeros(P + 1, dtype=np.int64)
    acc = 0
    for d in range(P):
        nb, rb = int(row_bytes[d]), recv_addr[d]
        lines = ((rb + nb + 15) >> 4) - ((rb >> 7) << 3) if nb else 0
        off[d] = acc
        acc += (lines + C - 1) // C
    off[P] = acc
    return off


class Memory:
    def __init__(self, send, send_addr, recv_addr, row_bytes):
        self.send, self.send_addr = send, send_addr
        self.recv_addr, self.row_bytes = recv_addr, row_bytes
        self.writes = {}

    def src(self, a):
        for s, base in enumerate(self.send_addr):
            if base <= a < base + len(self.send[s]):
                return int(self.send[s][a - base])
        raise AssertionError(f"read outside the send buffers at {a}")

    def write(self, a, v, d):
        lo = self.recv_addr[d]
        assert lo <= a < lo + self.row_bytes[d], f"write outside row {d} at {a}"
        assert a not in self.writes, f"byte {a} written twice"
        self.writes[a] = v


def search_le(a, n, x):
    """Largest i < n with a[i] <= x (wave_search_le)."""
    i = 0
    for j in range(n):
        if a[j] <= x:
            i = j
    return i


def seg_global(seg_end, seg_src, d, rbase, q):
    """a2av_seg_global: the segment holding row byte q (first end > q)."""
    se = seg_end[d]
    k = int(np.searchsorted(se, q, side="right"))
    st = int(se[k - 1]) if k > 0 else 0
    return st, int(se[k]), int(seg_src[d, k]) - (rbase + st)


def copy_kernel(P, seg_end, seg_src, row_bytes, recv_addr, chunk_off, C, U, defer, mem):
    """a2av_copy_kernel<U, NT, defer> over every chunk."""
    for ch in range(int(chunk_off[P])):
        d = search_le(chunk_off, P, ch)
        c = ch - int(chunk_off[d])
        rbase, rbytes = recv_addr[d], int(row_bytes[d])
        lend = (rbase + rbytes + 15) >> 4
        line0 = ((rbase >> 7) << 3) + c * C
        line1 = min(line0 + C, lend)
        assert line0 < line1, "empty chunk"
        se = seg_end[d]
        pos = line0
        while pos < line1:
            # window: 64 segments from the one holding the row offset of line `pos`
            o = max((pos << 4) - rbase, 0)
            s0 = search_le(se, P, o) + 1 if se[0] <= o else 0
            win = []
            for s in range(s0, min(s0 + WINDOW, P)):
                st = int(se[s - 1]) if s > 0 else 0
                en = int(se[s])
                if en > st:
                    win.append((st, en, int(seg_src[d, s]) - (rbase + st)))
            cover = rbytes if s0 + WINDOW >= P else int(se[s0 + WINDOW - 1])
            pend = min((rbase + cover) >> 4, line1)
            slow = pend <= pos
            if slow:
                pend = pos + 1
            lo, cov = (0, 0) if slow else (o, cover)
            sts = [w[0] for w in win]
            ens = [w[1] for w in win]
            Ds = [w[2] for w in win]
            cnt = len(win)
            pend_off = (pend << 4) - rbase
            nsearch = max(1, sum(1 for w in win if w[0] < pend_off))

            def lds(q):  # a2av_seg_lds
                k = 0
                for i in range(nsearch):
                    if sts[i] <= q:
                        k = i
                return k

            def seg_at(q):  # a2av_seg_at
                if lo <= q < cov:
                    k = lds(q)
                    assert win[k][0] <= q < win[k][1]
                    return win[k]
                return seg_global(seg_end, seg_src, d, rbase, q)

            def boundary(line):  # a2av_boundary_line<defer>
                la = line << 4
                off = la - rbase
                if defer and off >= 0 and off + 16 <= rbytes:
                    return
                for i in range(4):
                    q = off + 4 * i
                    if q >= 0 and q + 4 <= rbytes:
                        st, en, D = seg_at(q)
                        if q + 4 <= en and (la + 4 * i + D) % 4 == 0:
                            for b in range(4):
                                mem.write(la + 4 * i + b, mem.src(la + 4 * i + b + D), d)
                            continue
                    for j in range(4):
                        b = q + j
                        if 0 <= b < rbytes:
                            st, en, D = seg_at(b)
                            mem.write(la + 4 * i + j, mem.src(la + 4 * i + j + D), d)

            seen = []
            for base in range(pos, pend, 256 * U):
                for wv in range(4):
                    wb = base + wv * 64 * U
                    if wb >= pend:
                        continue
                    kc = lds(max((wb << 4) - rbase, lo)) if not slow else 0
                    # entries kc and kc + 1 cached per wave (bounds and source offsets)
                    kc1 = kc + 1 if kc + 1 < cnt else kc
                    Ekc, Dkc = ens[kc], Ds[kc]
                    Ekc1, Dkc1 = ens[kc1], Ds[kc1]
                    for u in range(U):
                        sl = wb + u * 64
                        if sl >= pend:
                            continue
                        soff = (sl << 4) - rbase
                        if soff >= Ekc and kc + 1 < nsearch:
                            if kc + 3 < nsearch and soff >= ens[kc + 3]:
                                kc = lds(soff)
                            else:
                                while kc + 1 < nsearch and soff >= ens[kc]:
                                    kc += 1
                            kc1 = kc + 1 if kc + 1 < cnt else kc
                            Ekc, Dkc = ens[kc], Ds[kc]
                            Ekc1, Dkc1 = ens[kc1], Ds[kc1]
                        one = soff + 16 * 64 <= Ekc
                        two = not one and not defer and soff + 16 * 64 <= Ekc1
                        for l in range(64):
                            line = sl + l
                            if line >= pend:
                                continue
                            seen.append(line)
                            off = soff + 16 * l
                            la = line << 4
                            if one:
                                # the slice lies in entry kc: one source offset for every lane
                                if off >= lo:
                                    assert win[kc][0] <= off and off + 16 <= win[kc][1]
                                    for b in range(16):
                                        mem.write(la + b, mem.src(la + b + Dkc), d)
                                else:
                                    boundary(line)
                                continue
                            if two:
                                # the slice spans entries kc and kc + 1
                                if off < lo:
                                    boundary(line)
                                elif off + 16 <= Ekc or off >= Ekc:
                                    Dl = Dkc if off < Ekc else Dkc1
                                    k = kc if off < Ekc else kc1
                                    assert win[k][0] <= off and off + 16 <= win[k][1]
                                    for b in range(16):
                                        mem.write(la + b, mem.src(la + b + Dl), d)
                                elif (Ekc - off) % 4 == 0 and (la + Dkc) % 4 == 0 and (la + Dkc1) % 4 == 0:
                                    for i in range(4):
                                        Dq = Dkc if off + 4 * i < Ekc else Dkc1
                                        for b in range(4):
                                            mem.write(la + 4 * i + b, mem.src(la + 4 * i + b + Dq), d)
                                else:
                                    boundary(line)
                                continue
                            if not (off >= lo and off + 16 <= cov):
                                boundary(line)
                                continue
                            kb = kc
                            if kc + 3 < cnt and off >= ens[kc + 3]:
                                kb = lds(off)
                            ks = [kb]
                            for _ in range(3):
                                ks.append(ks[-1] + 1 if ks[-1] + 1 < cnt else ks[-1])
                            e = [ens[k] for k in ks]
                            D = [Ds[k] for k in ks]
                            j = int(off >= e[0]) + int(off >= e[1]) + int(off >= e[2])
                            if off + 16 <= e[j]:
                                for b in range(16):
                                    mem.write(la + b, mem.src(la + b + D[j]), d)
                                continue
                            if defer:
                                boundary(line)
                                continue
                            addrs, ok = [], True
                            for i in range(4):
                                q = off + 4 * i
                                ji = int(q >= e[0]) + int(q >= e[1]) + int(q >= e[2])
                                a = la + 4 * i + D[ji]
                                ok = ok and a % 4 == 0 and q + 4 <= e[ji]
                                addrs.append(a)
                            if ok:
                                for i, a in enumerate(addrs):
                                    for b in range(4):
                                        mem.write(la + 4 * i + b, mem.src(a + b), d)
                            else:
                                boundary(line)
            assert sorted(seen) == list(range(pos, pend)), "a pass does not cover its lines once"
            pos = pend


def straddle_threads(P, seg_end, seg_src, row_bytes, recv_addr, mem):
    """a2av_straddle_line for every (row, segment)."""
    for d in range(P):
        se = seg_end[d]
        rbase, rbytes = recv_addr[d], int(row_bytes[d])
        for s in range(P):
            st = int(se[s - 1]) if s > 0 else 0
            en = int(se[s])
            if en <= st or (rbase + en) % 16 == 0:
                continue
            la = ((rbase + en) >> 4) << 4
            off = la - rbase
            if st > off or off < 0 or off + 16 > rbytes:
                continue
            kb = s + 1
            while se[kb] <= en:
                kb += 1
            eb = int(se[kb])
            a0 = int(seg_src[d, s]) + (off - st)
            a1 = int(seg_src[d, kb]) - en
            if eb >= off + 16:
                for i in range(4):
                    q = off + 4 * i
                    first = q < en
                    addr = a0 + 4 * i if first else a1 + q
                    mem.src(addr)  # the aligned-down dword read holds this source byte
                    if ((not first) or q + 4 <= en) and addr % 4 == 0:
                        for b in range(4):
                            mem.write(la + 4 * i + b, mem.src(addr + b), d)
                    else:
                        for j in range(4):
                            b = off + 4 * i + j
                            mem.write(la + 4 * i + j, mem.src(a0 + 4 * i + j if b < en else a1 + b), d)
                continue
            k = s
            for i in range(4):
                for j in range(4):
                    b = off + 4 * i + j
                    while b >= se[k]:
                        k += 1
                    sk = int(se[k - 1]) if k > 0 else 0
                    mem.write(la + 4 * i + j, mem.src(int(seg_src[d, k]) + (b - sk)), d)


def random_layout(rng, P, elem):
    mode = int(rng.integers(0, 4))
    if mode == 0:
        counts = rng.integers(0, 40, size=(P, P))
    elif mode == 1:
        counts = (rng.random((P, P)) < 0.1) * rng.integers(1, 300, size=(P, P))
    elif mode == 2:
        counts = rng.integers(0, 3, size=(P, P))
    else:
        counts = np.zeros((P, P), dtype=np.int64)
        counts[:, 0] = rng.integers(0, 500, size=P)
    sdisp = np.zeros_like(counts)
    sdisp[:, 1:] = np.cumsum(counts, axis=1)[:, :-1]
    send_bytes = counts.sum(axis=1) * elem + 8
    send = [rng.integers(0, 256, size=int(b), dtype=np.uint8) for b in send_bytes]
    base, send_addr = 1 << 20, []
    for s in range(P):
        base += int(rng.integers(0, 64)) * (4 if elem % 4 == 0 and rng.random() < 0.7 else 1)
        send_addr.append(base)
        base += int(send_bytes[s]) + 64
    rtot = counts.sum(axis=0) * elem
    adjacent = rng.random() < 0.5  # rows back to back, sharing lines, as the bench lays them out
    base = (base + 4096) // 16 * 16 + int(rng.integers(0, 16))
    recv_addr = []
    for d in range(P):
        recv_addr.append(base)
        base += int(rtot[d]) + (0 if adjacent else int(rng.integers(0, 40)))
    return counts, sdisp, send, send_addr, recv_addr


@pytest.mark.parametrize("defer", [0, 1])
@pytest.mark.parametrize("seed", range(16))
def test_every_row_byte_written_once(seed, defer):
    rng = np.random.default_rng(1000 * defer + seed)
    P = int(rng.choice([1, 2, 3, 5, 8, 17, 70, 130]))
    elem = int(rng.choice([1, 2, 4, 8]))
    U = int(rng.choice([1, 2]))
    C = 256 * U * int(rng.choice([1, 2])) if rng.random() < 0.3 else int(rng.choice([1, 2, 4, 8]))
    counts, sdisp, send, send_addr, recv_addr = random_layout(rng, P, elem)
    seg_end, seg_src, row_bytes = plan(P, elem, counts, sdisp, send_addr)
    chunk_off = scan(P, row_bytes, recv_addr, C)
    mem = Memory(send, send_addr, recv_addr, row_bytes)
    copy_kernel(P, seg_end, seg_src, row_bytes, recv_addr, chunk_off, C, U, defer, mem)
    if defer:
        straddle_threads(P, seg_end, seg_src, row_bytes, recv_addr, mem)
    assert len(mem.writes) == int(row_bytes.sum())
    for d in range(P):
        exp = b"".join(bytes(send[s][sdisp[s, d] * elem:(sdisp[s, d] + counts[s, d]) * elem])
                       for s in range(P))
        got = bytes(mem.writes[recv_addr[d] + i] for i in range(int(row_bytes[d])))
        assert got == exp, f"row {d} differs"


def test_model_catches_a_missing_owner():
    """Deferred copy without its straddle threads leaves multi-segment lines unwritten."""
    rng = np.random.default_rng(7)
    P, elem, C, U = 8, 4, 8, 1
    counts = np.full((P, P), 5, dtype=np.int64)  # 20-byte segments: most lines hold two
    sdisp = np.zeros_like(counts)
    sdisp[:, 1:] = np.cumsum(counts, axis=1)[:, :-1]
    send = [rng.integers(0, 256, size=int(counts[s].sum()) * elem, dtype=np.uint8) for s in range(P)]
    send_addr = [(1 << 20) + 4096 * s for s in range(P)]
    recv_addr = [(1 << 24) + 1024 * d + 4 for d in range(P)]
    seg_end, seg_src, row_bytes = plan(P, elem, counts, sdisp, send_addr)
    chunk_off = scan(P, row_bytes, recv_addr, C)
    mem = Memory(send, send_addr, recv_addr, row_bytes)
    copy_kernel(P, seg_end, seg_src, row_bytes, recv_addr, chunk_off, C, U, 1, mem)
    assert len(mem.writes) < int(row_bytes.sum())
    straddle_threads(P, seg_end, seg_src, row_bytes, recv_addr, mem)
    assert len(mem.writes) == int(row_bytes.sum())
