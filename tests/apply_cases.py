"""Hand-built change lists for apply_changes (bake/globe/erosion/particle.py:
821-1042), each with the outcome worked out by hand from the rules that
function states (its docstring and body), not by running any implementation.

Every case lives on one 4 x 4 face (16 cells, flat index c = i * 4 + j), all
cells active and at height 1 unless the case says otherwise, base level 0,
no sediment, pending or acc.  Amounts are dyadic, so the expected values are
exact in float64 and the comparison is bitwise.

Each case exercises one rule, chosen so that a plausible mistake in it (a
budget without ``acc``, the floor taken from the wrong neighbour, the
deficit not cancelling a deposit, the fan ceiling's sign, the write-off's
dust floor, the ring's next-older ceiling, a track written for a seafloor
step) changes an expected number.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

NE = 4
NCELL = NE * NE
NS = 16
S_SURF, S_ROUTE, S_SED, S_RFLAG, S_BASE = 0, 1, 7, 14, 15
DEFAULTS = dict(iter_erode=1.0, iter_deposit=4.0, fan_slope=0.125, dep_floor=0.0625, offshore_writeoff=0.25,
                use_route=0, use_erode_cap=0)


@dataclass
class Case:
    name: str
    droplets: list                       # per droplet: [(cell, delta, vol, (mom_a, mom_b)), ...]
    out: tuple                           # (clamped, to_pending, lost, lost_offshore)
    acc: dict                            # expected acc per cell (others 0)
    prm: dict = field(default_factory=dict)
    height: dict = field(default_factory=dict)
    sediment: dict = field(default_factory=dict)
    inactive: tuple = ()
    route: dict = field(default_factory=dict)   # cell -> (route surface, route flag)
    ecap: dict = field(default_factory=dict)
    pending: dict = field(default_factory=dict)
    disch: dict = field(default_factory=dict)
    mom: dict = field(default_factory=dict)
    rflag: dict = field(default_factory=dict)   # expected route flag after the apply

    def params(self):
        p = dict(DEFAULTS)
        p.update(self.prm)
        return p


M0 = (0.0, 0.0)
CASES = [
    # the budget iter_erode + acc[c]; the shortfall is a deficit that
    # outlives the list, i.e. lost; the sample's sediment channel clamps at 0
    Case("budget", [[(5, -0.5, 1.0, (0.5, 0.25))]], out=(1, 0.0, 0.375, 0.0), acc={5: -0.125},
         prm=dict(iter_erode=0.125), height={5: 5.0}, sediment={5: 0.0625}, disch={5: 1.0}, mom={5: (0.5, 0.25)}),
    # the budget is net per iteration: the second droplet sees the first's acc
    Case("budget_in_order", [[(5, -0.25, 1.0, M0)], [(5, -0.25, 0.5, M0)]], out=(1, 0.0, 0.125, 0.0),
         acc={5: -0.375}, prm=dict(iter_erode=0.375), height={5: 5.0}, disch={5: 1.5}),
    # floor = live surface of the next cell (5 -> 4.75 leaves 0.25); the
    # deficit 0.25 cancels the next deposit (0.375 -> 0.125), which the step
    # ceiling (cell 5's live surface 4.75 + dep_floor) cuts to 0.0625; the
    # surplus 0.0625 joins the final deposit
    Case("floor_deficit_ceiling", [[(5, -0.5, 1.0, M0), (6, 0.375, 0.75, M0), (7, 0.0625, -1.0, M0)]],
         out=(2, 0.0, 0.0, 0.0), acc={5: -0.25, 6: 0.0625, 7: 0.125}, height={5: 5.0, 6: 4.75, 7: 4.5},
         disch={5: 1.0, 6: 0.75}),
    # iter_deposit - acc[c] bounds each deposit; what the final deposit
    # cannot place goes to the (land) death cell's pending in full
    Case("land_surplus_to_pending", [[(9, 0.5, 1.0, M0), (10, 0.5, -1.0, M0)]], out=(2, 0.25, 0.0, 0.0),
         acc={9: 0.375, 10: 0.375}, prm=dict(iter_deposit=0.375), height={9: 3.0, 10: 3.0}, pending={10: 0.25},
         disch={9: 1.0}),
    # a submerged cell fills to -dep_floor at most; a submerged death cell
    # parks (1 - w) of the surplus and writes w off; a kept share under 1e-4
    # is dust, written off whole
    Case("sea_writeoff_and_dust", [[(13, 2.5, -1.0, M0)], [(14, 0.0625 + 2.0 ** -14, -1.0, M0)]],
         out=(2, 0.421875, 0.0, 0.140625 + 2.0 ** -14), acc={13: 1.9375, 14: 0.0625},
         height={13: -2.0, 14: -0.125}, pending={13: 0.421875}),
    # seafloor steps (volume 0): the second is held to the previous path
    # cell's live surface less fan_slope (-0.5 - 0.125), no discharge track;
    # with no final deposit the surplus leaves with the droplet
    Case("seafloor_fan", [[(13, 0.5, 0.0, M0), (14, 0.5, 0.0, M0)]], out=(1, 0.0, 0.0, 0.0),
         acc={13: 0.5, 14: 0.375}, height={13: -1.0, 14: -1.0}),
    # an inactive cell's entry is skipped, but it is still the next cell of
    # the entry before it (floor 4.75)
    Case("inactive_skipped", [[(5, -0.5, 1.0, M0), (0, -0.25, 1.0, M0)]], out=(1, 0.0, 0.25, 0.0),
         acc={5: -0.25}, height={0: 4.75, 5: 5.0}, inactive=(0,), disch={5: 1.0}),
    # a submerged cell is never eroded (floor = its own surface)
    Case("submerged_not_eroded", [[(13, -0.25, 1.0, M0)]], out=(1, 0.0, 0.25, 0.0), acc={},
         height={13: -1.0}, disch={13: 1.0}),
    # the final-deposit ring (death cell 7, then the path back through 6
    # and 5): each ring entry is held to the NEXT-OLDER ring cell's live
    # surface + dep_floor -- 7 to 6's 4.375 (0.4375 placed), 6 to 5's 4.875
    # (0.5625), not to the last step cell -- and the oldest has no ceiling
    Case("final_ring", [[(5, -0.125, 1.0, M0), (6, -0.125, 1.0, M0), (7, 2.0, -1.0, M0), (6, 0.0, -1.0, M0),
                         (5, 0.0, -1.0, M0)]], out=(2, 0.0, 0.0, 0.0),
         acc={5: 0.875, 6: 0.4375, 7: 0.4375}, height={5: 5.0, 6: 4.5, 7: 4.0}, disch={5: 1.0, 6: 1.0}),
    # with the route: a lake cell (flag 2) is not eroded below its water
    # level; an ordinary cell's flag follows route != live surface
    Case("route_floor_and_flags", [[(5, -0.5, 1.0, M0), (6, -0.125, 1.0, M0)], [(7, 0.25, -1.0, M0)]],
         out=(1, 0.0, 0.375, 0.0), acc={5: -0.125, 6: -0.125, 7: 0.25}, prm=dict(use_route=1),
         height={5: 5.0, 6: 4.75, 7: 4.5}, route={5: (4.875, 2.0), 6: (4.75, 0.0), 7: (4.75, 1.0)},
         disch={5: 1.0, 6: 1.0}, rflag={5: 2.0, 6: 1.0, 7: 0.0}),
    # the slope limit: the budget is also at most erode_cap[c] + acc[c]
    Case("erode_cap", [[(5, -0.25, 1.0, M0)]], out=(1, 0.0, 0.1875, 0.0), acc={5: -0.0625},
         prm=dict(use_erode_cap=1), height={5: 5.0}, ecap={5: 0.0625}, disch={5: 1.0}),
]


def inputs(case: Case):
    """numpy inputs of one case: (lists, fields, cap) with lists =
    (cl_cell, cl_delta, cl_vol, cl_mom, cl_count) and fields = dict of the
    state arrays, shaped (1, NE, NE[, k])."""
    P = len(case.droplets)
    cap = max(len(d) for d in case.droplets)
    cl_cell = np.zeros(P * cap, np.int32)
    cl_delta = np.zeros(P * cap, np.float64)
    cl_vol = np.zeros(P * cap, np.float32)
    cl_mom = np.zeros((P * cap, 2), np.float32)
    cl_count = np.zeros(P, np.int64)
    for p, d in enumerate(case.droplets):
        cl_count[p] = len(d)
        for k, (c, dl, v, m) in enumerate(d):
            cl_cell[p * cap + k], cl_delta[p * cap + k], cl_vol[p * cap + k] = c, dl, v
            cl_mom[p * cap + k] = m
    h = np.ones(NCELL)
    s = np.zeros(NCELL)
    for c, v in case.height.items():
        h[c] = v
    for c, v in case.sediment.items():
        s[c] = v
    mask = np.ones(NCELL, np.uint8)
    mask[list(case.inactive)] = 0
    samp = np.zeros((NCELL, NS), np.float32)
    samp[:, S_SURF] = (h + s).astype(np.float32)
    samp[:, S_SED] = s.astype(np.float32)
    samp[:, S_ROUTE] = samp[:, S_SURF]
    for c, (r, fl) in case.route.items():
        samp[c, S_ROUTE], samp[c, S_RFLAG] = r, fl
    ecap = np.zeros(NCELL)
    for c, v in case.ecap.items():
        ecap[c] = v
    sh = (1, NE, NE)
    fields = dict(height=h.reshape(sh), sediment=s.reshape(sh), acc=np.zeros(sh), pending=np.zeros(sh),
                  samp=samp.reshape(sh + (NS,)), disch_track=np.zeros(sh), mom_track=np.zeros(sh + (2,)),
                  mask=mask.reshape(sh), erode_cap=ecap.reshape(sh))
    return (cl_cell, cl_delta, cl_vol, cl_mom, cl_count), fields, cap


def check(case: Case, fields: dict, out4) -> None:
    """Assert the applied state equals the hand-derived outcome, bitwise."""
    dense = lambda d, shape=(NCELL,): _dense(d, shape)
    assert tuple(float(x) for x in out4) == tuple(float(x) for x in case.out), (case.name, tuple(out4), case.out)
    acc = np.asarray(fields["acc"]).reshape(NCELL)
    assert np.array_equal(acc, dense(case.acc)), (case.name, acc)
    assert np.array_equal(np.asarray(fields["pending"]).reshape(NCELL), dense(case.pending)), case.name
    assert np.array_equal(np.asarray(fields["disch_track"]).reshape(NCELL), dense(case.disch)), case.name
    mom = np.asarray(fields["mom_track"]).reshape(NCELL, 2)
    assert np.array_equal(mom, dense(case.mom, (NCELL, 2))), (case.name, mom)
    samp = np.asarray(fields["samp"]).reshape(NCELL, NS)
    h, s = np.ones(NCELL), np.zeros(NCELL)
    for c, v in case.height.items():
        h[c] = v
    for c, v in case.sediment.items():
        s[c] = v
    # the packed samples follow the live surface on every active cell an entry names
    touched = sorted({c for d in case.droplets for (c, *_r) in d} - set(case.inactive))
    for c in touched:
        assert samp[c, S_SURF] == np.float32(h[c] + s[c] + acc[c]), (case.name, c)
        assert samp[c, S_SED] == np.float32(max(s[c] + acc[c], 0.0)), (case.name, c)
    for c, v in case.rflag.items():
        assert samp[c, S_RFLAG] == np.float32(v), (case.name, c, samp[c, S_RFLAG])


def _dense(d: dict, shape):
    a = np.zeros(shape)
    for c, v in d.items():
        a[c] = v
    return a
