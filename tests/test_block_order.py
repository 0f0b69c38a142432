"""The workgroup renumberings of the flat launches (xcd_block and run_block
in bonzo_amd/csrc/bnz_common.hpp), restated with the device's integer
arithmetic: each must be a permutation of the launch's workgroups (every
cell is then computed exactly once), and must give XCD x = b % 8 the
logical blocks the comment promises."""
NXCD = 8
XCD_RUN = 16


def xcd_block(b, nb):
    q, r = nb // NXCD, nb % NXCD
    x, idx = b % NXCD, b // NXCD
    return x * q + (x if x < r else r) + idx


def run_block(b, nb):
    nfull = nb // (NXCD * XCD_RUN) * (NXCD * XCD_RUN)
    if b >= nfull:
        return b
    x, idx = b % NXCD, b // NXCD
    return ((idx // XCD_RUN) * NXCD + x) * XCD_RUN + idx % XCD_RUN


SIZES = [1, 2, 7, 8, 9, 127, 128, 129, 255, 256, 1000, 4095, 4096, 4097, 65536, 67081, 70313]


def test_both_orders_are_permutations():
    for nb in SIZES:
        for f in (xcd_block, run_block):
            assert sorted(f(b, nb) for b in range(nb)) == list(range(nb)), (f.__name__, nb)


def test_xcd_block_gives_each_xcd_one_contiguous_run():
    for nb in SIZES:
        for x in range(NXCD):
            mine = sorted(xcd_block(b, nb) for b in range(x, nb, NXCD))
            assert mine == list(range(mine[0], mine[0] + len(mine))) if mine else True


def test_run_block_deals_runs_of_16_round_robin():
    for nb in SIZES:
        nfull = nb // (NXCD * XCD_RUN) * (NXCD * XCD_RUN)
        for b in range(nfull):
            lb = run_block(b, nb)
            assert (lb // XCD_RUN) % NXCD == b % NXCD  # logical run r runs on XCD r % 8
