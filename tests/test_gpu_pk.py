"""Persistent-kernel eager hash join against the oracle.

The kernel's bounded idle wait is forced from the host (SJB_PK_IDLE_POLLS)
so that it exits between batches and the relaunch path is exercised; every
kernel still exits normally.
"""
import os

import pytest

from helpers import make

pytestmark = pytest.mark.gpu


def run_windows(sjbmod, orc, B=2048, nb=6, rounds=2, idle=None):
    if idle is not None:
        os.environ["SJB_PK_IDLE_POLLS"] = str(idle)
    try:
        W = B * nb
        R, S = make(W, 3000, 141), make(W, 3000, 142)
        e = orc.Eager(1000, W)
        with sjbmod.Context(max_window=W, batch_size=B, flags=1) as c:
            for _ in range(rounds):
                for b in range(nb):
                    for is_left, src in ((1, R), (0, S)):
                        batch = src[b * B:(b + 1) * B]
                        rc, st = c.pk_execute(batch, is_left)
                        assert rc == 0, rc
                        o = e.execute(batch, is_left)
                        assert (st.result_count, st.checksum1, st.checksum2) == o.triple()
                c.pk_clear()
                e.clear()
            launches = c.pk_launches()
            # a second batch larger than a slot and a full side are rejected
            rc, _ = c.pk_execute(make(B + 1, 10, 1), 1)
            assert rc == -2
            # the eager (per-launch) path may not share the persistent state
            rc, _ = c.eager_execute(2, 0, 0, R[:B], 1)
            assert rc == -9
            c.pk_stop()
        return launches
    finally:
        os.environ.pop("SJB_PK_IDLE_POLLS", None)


def test_pk_matches_oracle(sjbmod, orc):
    print("case pk default", flush=True)
    run_windows(sjbmod, orc)


def test_pk_relaunch_after_idle_exit(sjbmod, orc):
    print("case pk idle relaunch", flush=True)
    launches = run_windows(sjbmod, orc, nb=3, rounds=2, idle=4)
    assert launches > 1
