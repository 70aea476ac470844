"""Writes patterns.bin, patterns.json and trdump.bin for the LDS probe.

Two kinds of entries, one probe dispatch each:

- pair tests: only lanes a and b active, reading or writing addresses 0 and
  2048 (the same banks, different words). A conflict cycle shows that the two
  lanes are served in the same pass, so the list of b that conflict with a
  gives the lane grouping of each access width;
- the LDS accesses of the denoiser's kernels, as they are laid out now and
  as they were before (the MFMA operand reads of k_conv and k_conv_pipe, the
  weight gradient's operand reads and its staging writes), every read of one
  chunk or tile in one dispatch.
"""
import json
import struct

NA = 0xFFFFFFFF
WIDTH = {0: 16, 1: 8, 2: 2, 3: 8, 4: 8, 5: 16}
KIND = {0: "ds_read_b128", 1: "ds_write_b64", 2: "ds_write_b16", 3: "ds_read_b64", 4: "ds_read_b64_tr_b16",
        5: "ds_write_b128"}
HR, HC = 6, 66          # halo rows and columns of a 4 x 64 tile
WR, WC, WX = 4, 64, 80  # weight-gradient band rows, tile columns, staged x columns

entries = []


def add(name, kind, pats):
    for p in pats:
        assert len(p) == 64, name
        for a in p:
            assert a == NA or (a % WIDTH[kind] == 0 and 0 <= a <= 65536 - WIDTH[kind]), (name, a)
    entries.append((name, kind, pats))


def pair_tests():
    for kind in range(6):
        add(f"{KIND[kind]} contiguous", kind, [[WIDTH[kind] * l for l in range(64)]])
        for a in (0, 4, 16, 32):
            add(f"{KIND[kind]} single a={a}", kind, [[0 if l == a else NA for l in range(64)]])
            for b in range(64):
                if b != a:
                    add(f"{KIND[kind]} pair a={a} b={b}", kind,
                        [[0 if l == a else 2048 if l == b else NA for l in range(64)]])


def operand_reads(G, NT, mirror, row_bytes, swz, weights_base):
    """The 16-byte MFMA operand reads of one chunk of G 16-channel groups:
    halo reads of the 4 waves and weight reads of wave 0, piece q of pixel
    (weight row) p at byte p * row_bytes + 16 * (q ^ swz(p))."""
    halo, wts = [], []
    for wave in range(4):
        for i in range((9 * G + 1) // 2):
            hp = [[0] * 64 for _ in range(4)]
            wp = [[0] * 64 for _ in range(NT)]
            for lane in range(64):
                ln = lane if 2 * i + (lane >> 5) < 9 * G else lane - 32  # past the end: a copy of the first half
                sub = 2 * i + (ln >> 5)
                half, col = (ln >> 4) & 1, ln & 15
                tap, q = sub // G, (sub % G) * 2 + half
                dy, dx = tap // 3 - 1, tap % 3 - 1
                if mirror:
                    dy, dx = -dy, -dx
                for f in range(4):
                    p = (wave + 1 + dy) * HC + col + 1 + dx + 16 * f
                    hp[f][lane] = p * row_bytes + 16 * (q ^ swz(p))
                for j in range(NT):
                    r = tap * NT * 16 + j * 16 + col
                    wp[j][lane] = weights_base + r * row_bytes + 16 * (q ^ swz(r))
            halo += hp
            if wave == 0:
                wts += wp
    return halo, wts


def kernel_layouts():
    none = lambda p: 0
    layouts = [
        ("k_conv_pipe 16-channel chunks, now: 32-byte pixels", 1, 32, none),
        ("k_conv_pipe 16-channel chunks, before: pieces ^ (p >> 3) & 1", 1, 32, lambda p: (p >> 3) & 1),
        ("k_conv_pipe 32-channel chunks, now: pieces ^ (p >> 1) & 3", 2, 64, lambda p: (p >> 1) & 3),
        ("k_conv_pipe 32-channel chunks, before: pieces ^ (p >> 2) & 3", 2, 64, lambda p: (p >> 2) & 3),
        ("k_conv 32-channel chunks, now: 64-byte rows, pieces ^ (p >> 1) & 3", 2, 64, lambda p: (p >> 1) & 3),
        ("k_conv 16-channel chunks, now: 64-byte rows, pieces ^ (p >> 1) & 3", 1, 64, lambda p: (p >> 1) & 3),
        ("k_conv 32-channel chunks, before: 80-byte rows", 2, 80, none),
        ("k_conv 16-channel chunks, before: 80-byte rows", 1, 80, none),
    ]
    for name, G, rb, swz in layouts:
        halo, _ = operand_reads(G, 1, False, rb, swz, 0)
        add(f"{name}: halo reads", 0, halo)
        halo, _ = operand_reads(G, 1, True, rb, swz, 0)
        add(f"{name}: halo reads, mirrored taps", 0, halo)
        for NT in (1, 2, 3):
            _, wts = operand_reads(G, NT, False, rb, swz, HR * HC * rb % 1024)  # the banks of the real base
            add(f"{name}: weight reads, NT={NT}", 0, wts)


def wgrad_layouts():
    """The removed kernel k_wgrad: transposed staging and 16-byte operand reads."""
    for kSD, kSX, when in ((272, 496, "2 x 2 fragments"), (264, 488, "before")):
        dT0, xT0 = 0, 2 * 16 * kSD
        a_reads, x_reads = [], []
        for wave in range(4):
            for kc in range(2):
                for a in range(2):
                    a_reads.append([2 * (dT0 + (a * 16 + (l & 15)) * kSD + wave * WC + 32 * kc + 8 * (l >> 4))
                                    for l in range(64)])
                for b in range(2):
                    for dr in range(3):
                        x_reads.append([2 * (xT0 + (b * 16 + (l & 15)) * kSX + (wave + dr) * WX + 8 + 32 * kc +
                                             8 * (l >> 4)) for l in range(64)])
        add(f"k_wgrad (removed) dPre operand reads, strides {kSD}/{kSX}, {when}", 0, a_reads)
        add(f"k_wgrad (removed) x operand reads (centre piece), strides {kSD}/{kSX}, {when}", 0, x_reads)
    # staging writes
    for kS, DQ in ((264, 2), (264, 4)):
        pats = []
        for wave in range(4):
            for e in range(8):
                pats.append([2 * ((8 * ((64 * wave + l) % DQ) + e) * kS + ((64 * wave + l) // DQ) // 64 * 64 +
                                  ((64 * wave + l) // DQ) % 64) for l in range(64)])
        add(f"k_wgrad (removed) dPre staging, 2-byte stores, {DQ} pieces per pixel, stride {kS}", 2, pats)
    pats = []
    for wave in range(4):
        for e in range(8):
            p = []
            for l in range(64):
                i = 64 * wave + l
                q, g = i >> 6, i & 63
                p.append(2 * ((8 * q + e) * 272 + (g >> 4) * 64 + (g & 15) * 4))
            pats.append(p)
    add("k_wgrad (removed) dPre staging, 4 x 8 register transposes, 8-byte stores, stride 272", 1, pats)


def wgrad_tr_layouts():
    """k_wgrad_tr's operand reads (ds_read_b64_tr_b16): lane l of 16-lane group g names pixel
    p0 + 8 g + 4 h + (l >> 2) & 3 and 8-byte chunk l & 3 of a fragment region of 32-byte pixel rows,
    slots swizzled by bit 3 of the pixel (and, for comparison, not swizzled)."""
    swizzled = lambda p, k: (2 * p + k) ^ (((p >> 3) & 1) << 3)
    plain = lambda p, k: 2 * p + k
    for TC in (32, 64):
        for name, slot in (("swizzled by pixel bit 3", swizzled), ("plain", plain)):
            for what in ("dPre", "x"):
                pats = []
                taps = [(0, 0)] if what == "dPre" else [(dr, dx) for dr in range(3) for dx in (-1, 0, 1)]
                for wave in range(4):
                    for kc in range(TC // 32):
                        for h in range(2):
                            for dr, dx in taps:
                                p0 = wave * TC + 32 * kc if what == "dPre" else (wave + dr) * (TC + 2) + 32 * kc + 1 + dx
                                pats.append([slot(p0 + 8 * (l >> 4) + 4 * h + ((l >> 2) & 3), (l & 3) >> 1) * 16 +
                                             (l & 1) * 8 for l in range(64)])
                add(f"k_wgrad_tr {what} operand reads, {TC}-column tiles, {name}", 4, pats)


def main():
    pair_tests()
    kernel_layouts()
    wgrad_layouts()
    wgrad_tr_layouts()
    with open("patterns.bin", "wb") as fo:
        fo.write(struct.pack("<i", len(entries)))
        for _, kind, pats in entries:
            fo.write(struct.pack("<ii", kind, len(pats)))
            for p in pats:
                fo.write(struct.pack("<64I", *p))
    json.dump([[n, k, len(p)] for n, k, p in entries], open("patterns.json", "w"))
    # the transpose read: lane l names row (l >> 2) & 3 of 64-byte rows, 8-byte chunk l & 3, 16-lane group l >> 4
    tr = [((l >> 2) & 3) * 64 + (l & 3) * 8 + (l >> 4) * 512 for l in range(64)]
    with open("trdump.bin", "wb") as fo:
        fo.write(struct.pack("<64I", *tr))
    print(f"{len(entries)} dispatches")


if __name__ == "__main__":
    main()
