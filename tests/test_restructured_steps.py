"""Steps the device code restructures for fewer dependent memory accesses, checked by brute force against the literal
sequence they replace (which follows the reference line by line). The whole-tick parity tests cover them again in
context."""


def test_move_cells_equals_remove_then_push(oracle):
    """grid_broadphase.cuh:44-84: removing the old block's cells one by one, then pushing the new block's, equals the
    one-pass move_cells, including the overflow count when the list is full."""
    for p in ((0, 0), (6, 2)):
        w = oracle.build_world(*p)
        assert oracle.check_move_cells(w, 12345 + p[0], 100000) == 0
