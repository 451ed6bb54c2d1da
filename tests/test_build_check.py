"""The build's check on spill placement in the device assembly (breseq_amd/build.py): a VGPR
spill stored before a block's exec-mask restore runs under the narrowed mask of the branch that
block closes, so lanes outside the branch reload a stale slot."""
from breseq_amd import build

BAD = """_ZN3bsq1kEv:
.LBB0_7:
\tv_writelane_b32 v152, s38, 45
\tscratch_store_dwordx2 off, v[118:119], off offset:64 ; 8-byte Folded Spill
\ts_or_b64 exec, exec, s[20:21]
\ts_swappc_b64 s[30:31], s[0:1]
\tscratch_load_dwordx2 v[118:119], off, off offset:64 ; 8-byte Folded Reload
"""

GOOD = """_ZN3bsq1kEv:
.LBB0_7:
\ts_or_b64 exec, exec, s[20:21]
\tscratch_store_dwordx2 off, v[118:119], off offset:64 ; 8-byte Folded Spill
\ts_swappc_b64 s[30:31], s[0:1]
\tscratch_load_dwordx2 v[118:119], off, off offset:64 ; 8-byte Folded Reload
.LBB0_8:
\tscratch_store_dword off, v3, off offset:8 ; 4-byte Folded Spill
\ts_branch .LBB0_9
.LBB0_9:
\ts_or_b64 exec, exec, s[2:3]
"""


def test_flags_a_spill_ahead_of_the_exec_restore():
    hits = build.spill_before_exec_restore(BAD)
    assert hits == [("_ZN3bsq1kEv", 2, "scratch_store_dwordx2 off, v[118:119], off offset:64")]


def test_passes_spills_after_the_restore_or_in_other_blocks():
    assert build.spill_before_exec_restore(GOOD) == []
