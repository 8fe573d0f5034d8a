"""Recipes of the golden fixtures: system expression, batch, temperatures, seeds and sweep mode.

The same expression is evaluated against the reference package (make_golden.py) and against this build's
sck_amd.system (the tests), so both sides build the same system. `sy`, `cr` and `np` are the names in scope.
"""

MASKED_SQUARE = ("sy.System('Ising square masked h', 1, cr.make('square'), (8, 8), (True, True), {1: 1.0}, h=(0.3,), "
                 "mask=np.random.default_rng(0).random((1, 1, 8, 8)) > 0.25)")
MASKED_OPEN_SC = ("sy.System('Ising sc open masked', 1, cr.make('sc'), (4, 3, 3), (False,) * 3, {1: 1.0}, "
                  "mask=np.concatenate([np.ones((1, 1, 3, 4), bool), np.array([[[[1, 1, 1, 1], [1, 0, 0, 1], [1, 1, 1, 1]]]], bool), "
                  "np.ones((1, 1, 3, 4), bool)], axis=1))")

RECIPES = {
    "ising_square8": dict(system="sy.ising('square', 8)", B=4, T=[1.5, 2.0, 2.5, 3.0], seed=3, init_seed=1, sweeps=3,
                          sampler=dict(packed=False)),
    "ising_square8_barker": dict(system="sy.ising('square', 8)", acceptance="barker", B=4, T=[1.5, 2.0, 2.5, 3.0],
                                 seed=3, init_seed=1, sweeps=3, sampler=dict(packed=False)),
    "ising_square8_seq": dict(system="sy.ising('square', 4)", B=2, T=[2.2, 2.2], seed=5, init_seed=2, sweeps=2,
                              sampler=dict(packed=False), mode="reference"),
    "ising_masked_h": dict(system=MASKED_SQUARE, B=3, T=[1.0, 2.0, 3.0], seed=7, init_seed=4, sweeps=2,
                           sampler=dict(packed=False)),
    "ising_open_masked_sc": dict(system=MASKED_OPEN_SC, B=2, T=[2.5, 4.0], seed=9, init_seed=5, sweeps=2,
                                 sampler=dict(packed=False)),
    "ising_triangular6": dict(system="sy.ising('triangular', 6)", B=2, T=[3.0, 4.0], seed=2, init_seed=3, sweeps=2,
                              sampler=dict(packed=False)),
    "ising_j1j2_sc": dict(system="sy.System('J1-J2', 1, cr.make('sc', (1, 2)), (4, 4, 4), (True,) * 3, {1: 1.0, 2: 0.3})",
                          B=2, T=[2.0, 5.0], seed=4, init_seed=6, sweeps=2, sampler=dict(packed=False)),
    "ising_packed40": dict(system="sy.ising('square', 4)", B=40, T=[1.5 + 2.5 * i / 39 for i in range(40)], seed=5,
                           init_seed=9, sweeps=2),
    "xy_square4": dict(system="sy.xy('square', 4)", B=2, T=[0.8, 1.6], seed=11, init_seed=2, sweeps=2),
    "heis_bcc2": dict(system="sy.heisenberg('bcc', 2)", B=2, T=[1.5, 1.5], seed=11, init_seed=2, sweeps=2),
    "heis_kagome3": dict(system="sy.heisenberg('kagome', 3, J=-1.0)", B=2, T=[0.4, 1.0], seed=1, init_seed=3, sweeps=2),
    "nanodot_r2": dict(system="sy.nanodot(Rd=2.0, L=2)", B=2, T=[0.7, 1.4], seed=3, init_seed=0, sweeps=2),
    "nanodot_shells": dict(system="sy.nanodot(Rd=2.5, L=3, shells=(1, 2, 4), J_shells={2: 0.3, 4: -0.1})", B=2,
                           T=[0.5, 0.5], seed=3, init_seed=3, sweeps=1),
    "heis_sc_barker": dict(system="sy.heisenberg('sc', 4)", acceptance="barker", B=2, T=[1.0, 2.0], seed=8,
                           init_seed=1, sweeps=1),
    "sync_ising_square4": dict(system="sy.ising('square', 4)", B=2, T=[2.2, 2.2], seed=3, init_seed=1, sweeps=1,
                               sampler=dict(packed=False), mode="sync"),
    "sync_xy_square4": dict(system="sy.xy('square', 4)", B=2, T=[0.9, 0.9], seed=3, init_seed=1, sweeps=1, mode="sync"),
    "bitwise_square8": dict(system="sy.ising('square', 8)", B=64, T=[1.5] * 32 + [3.0] * 32, seed=13, init_seed=3,
                            sweeps=2, sampler=dict(bitwise=True)),
    "bitwise_square8_barker": dict(system="sy.ising('square', 8)", acceptance="barker", B=64, T=[1.5] * 32 + [3.0] * 32,
                                   seed=13, init_seed=3, sweeps=2, sampler=dict(bitwise=True)),
    "bitwise_indep_square8": dict(system="sy.ising('square', 8)", B=64, T=[1.5] * 32 + [3.0] * 32, seed=13,
                                  init_seed=3, sweeps=2, sampler=dict(bitwise=True, independent=True)),
    "bitwise_indep_masked_h": dict(system=MASKED_SQUARE, B=64, T=[1.0] * 32 + [2.5] * 32, seed=8, init_seed=6,
                                   sweeps=2, sampler=dict(bitwise=True, independent=True)),
    "bitwise_masked_h": dict(system=MASKED_SQUARE, B=64, T=[1.0] * 32 + [2.5] * 32, seed=8, init_seed=6, sweeps=2,
                             sampler=dict(bitwise=True)),
    "bitwise_indep_sc5_barker": dict(system="sy.System('Ising sc odd', 1, cr.make('sc'), (5, 5, 5), (True,) * 3, {1: 1.0})",
                                     acceptance="barker", B=96, T=[4.5] * 96, seed=2, init_seed=1, sweeps=2,
                                     sampler=dict(bitwise=True, independent=True)),
}
