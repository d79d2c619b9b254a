"""Test-side mirrors of the reference modules that drive the drop-in
(config.py, defaults.py, populations.py): the population objects, the flag
dictionary and the module globals the reference hands to
updates.SRK2_force_calc, so that the drop-in tests call it the way
control.sim_loop does (control.py:24-84)."""
