"""CPU oracle of the hydro hot path -- test infrastructure only (tests, smoke check, CPU baseline)."""
