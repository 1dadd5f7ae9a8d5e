"""Block structures of libxm's own test driver, restated as recipes.

Each builder returns Spec objects that replay, call for call, the structure
set-up of the corresponding libxm test.c function (cited per builder); the
tables at the bottom are libxm's contraction and addition test tables
(libxm test.c:1878-1959).
"""
from libxm_amd import xm
from libxm_amd.spec import Spec

D = xm.SCALAR_DOUBLE

# The 8 permutations of libxm's (ij)(ab)<->(ab)(ij) symmetric 4-index test
# tensors, in the order test.c applies them (libxm test.c:835-852, 872-892).
PERMS8 = [((1, 0, 2, 3), -1), ((0, 1, 3, 2), -1), ((1, 0, 3, 2), 1),
          ((2, 3, 0, 1), 1), ((3, 2, 0, 1), -1), ((2, 3, 1, 0), -1),
          ((3, 2, 1, 0), 1)]
PERMS4 = PERMS8[:3]


def _lex(nb):
    """Block indices with the LAST index fastest (nested C loops)."""
    out = [()]
    for n in nb:
        out = [t + (i,) for t in out for i in range(n)]
    return out


def _symmetric(specs, perms, guard=None):
    """libxm test.c loop: visit blocks last-index-fastest, make every still
    zero block canonical in all specs and set its images under `perms` as
    derivative blocks where they are still zero (checked in `guard`)."""
    guard = guard or specs[0]
    for idx in _lex(guard.nblocks):
        if guard.kind(idx) is not None:
            continue
        for s in specs:
            s.canon(idx)
        for perm, sc in perms:
            tgt = tuple(idx[p] for p in perm)
            if guard.kind(tgt) is None:
                for s in specs:
                    s.deriv(tgt, idx, perm, sc)


def abc_1(t=D):  # libxm test.c:523-542
    a, b, c = (Spec((4, 4), None, t) for _ in range(3))
    for s in (a, b, c):
        s.canon((0, 0))
    return a, b, c


def abc_2(t=D):  # libxm test.c:544-572
    sp = {0: [2], 1: [2]}
    return tuple(Spec((4, 4), sp, t).all_canonical() for _ in range(3))


def abc_3(t=D):  # libxm test.c:574-591
    a = Spec((1,) * 8, None, t).all_canonical()
    b = Spec((1,) * 8, None, t).all_canonical()
    c = Spec((1, 1), None, t).all_canonical()
    return a, b, c


def abc_4(t=D):  # libxm test.c:593-608
    return tuple(Spec((3,) * 6, None, t).all_canonical() for _ in range(3))


def abc_5(t=D):  # libxm test.c:610-641
    sp = {0: [2], 1: [2]}
    a, b, c = (Spec((4, 4), sp, t) for _ in range(3))
    for s in (a, b, c):
        s.canon((0, 0))
    c.canon((0, 1))
    c.canon((1, 0))
    for s in (a, b, c):
        s.canon((1, 1))
    return a, b, c


def abc_6(t=D):  # libxm test.c:643-677
    a = Spec((3,), {0: [1]}, t)
    b = Spec((2, 5), {1: [3]}, t)
    c = Spec((2, 3, 5), {1: [1], 2: [3]}, t)
    a.canon((0,)).canon((1,))
    b.canon((0, 0)).canon((0, 1))
    for idx in [(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1)]:
        c.canon(idx)
    return a, b, c


def abc_7(t=D):  # libxm test.c:679-738
    a = Spec((7, 4, 11), {0: [2, 5], 1: [2], 2: [2, 9]}, t)
    a.canon((0, 0, 0)).canon((1, 0, 0))
    a.deriv((2, 0, 0), (0, 0, 0), (2, 1, 0), -0.5)
    a.deriv((0, 1, 0), (0, 0, 0), (1, 0, 2), 1.5)
    a.canon((1, 1, 0))
    a.deriv((2, 1, 0), (0, 0, 0), (2, 0, 1), 0.7)
    a.all_canonical()
    b = Spec((7, 4, 8), {0: [2, 5], 1: [2], 2: [2, 6]}, t).all_canonical()
    c = Spec((8, 11), {0: [2, 6], 1: [2, 9]}, t).all_canonical()
    return a, b, c


def abc_8(t=D):  # libxm test.c:740-776
    a = Spec((15, 10), {0: [5, 10], 1: [5]}, t)
    a.canon((0, 0))
    a.deriv((0, 1), (0, 0), (1, 0), -0.3)
    a.canon((2, 0)).canon((2, 1))
    b = Spec((10,), {0: [5]}, t).canon((0,)).canon((1,))
    c = Spec((15,), {0: [5, 10]}, t).canon((0,)).canon((2,))
    return a, b, c


def abc_9(t=D):  # libxm test.c:778-896
    o_sp = [2, 5]
    v_sp = [3, 4, 8]
    a = Spec((6, 6, 10, 10), {0: o_sp, 1: o_sp, 2: v_sp, 3: v_sp}, t)
    c = Spec((6, 6, 10, 10), {0: o_sp, 1: o_sp, 2: v_sp, 3: v_sp}, t)
    b = Spec((10,) * 4, {k: v_sp for k in range(4)}, t)
    _symmetric([a, c], PERMS4)
    _symmetric([b], PERMS8)
    return a, b, c


def abc_10(t=D):  # libxm test.c:898-1000 (a and b swapped on return)
    o_sp, v_sp = [4], [3, 7]
    a = Spec((6, 6, 9, 9), {0: o_sp, 1: o_sp, 2: v_sp, 3: v_sp}, t)
    c = Spec((6, 6, 9, 9), {0: o_sp, 1: o_sp, 2: v_sp, 3: v_sp}, t)
    _symmetric([a, c], PERMS4)
    b = Spec((9,) * 4, {k: v_sp for k in range(4)}, t)
    _symmetric([b], PERMS8)
    return b, a, c


def abc_11(t=D):  # libxm test.c:1002-1087
    a = Spec((6,) * 4, {k: [2] for k in range(4)}, t)
    a.canon((0, 0, 0, 0)).canon((1, 0, 0, 0))
    a.deriv((0, 1, 0, 0), (1, 0, 0, 0), (1, 0, 2, 3), -1)
    a.canon((1, 1, 0, 0))
    a.deriv((0, 0, 1, 0), (1, 0, 0, 0), (2, 3, 0, 1), 1)
    a.canon((1, 0, 1, 0))
    a.deriv((0, 1, 1, 0), (1, 0, 1, 0), (1, 0, 2, 3), -1)
    a.canon((1, 1, 1, 0))
    a.deriv((0, 0, 0, 1), (1, 0, 0, 0), (2, 3, 1, 0), -1)
    a.deriv((1, 0, 0, 1), (1, 0, 1, 0), (0, 1, 3, 2), -1)
    a.deriv((0, 1, 0, 1), (1, 0, 1, 0), (1, 0, 3, 2), 1)
    a.deriv((1, 1, 0, 1), (1, 1, 1, 0), (0, 1, 3, 2), -1)
    a.deriv((0, 0, 1, 1), (1, 1, 0, 0), (2, 3, 0, 1), 1)
    a.deriv((1, 0, 1, 1), (1, 1, 1, 0), (2, 3, 0, 1), 1)
    a.deriv((0, 1, 1, 1), (1, 1, 1, 0), (3, 2, 0, 1), -1)
    a.canon((1, 1, 1, 1))
    b = Spec((3, 3, 6, 6), {2: [2], 3: [2]}, t)
    for idx in [(0, 0, 0, 0), (0, 0, 0, 1), (0, 0, 1, 0), (0, 0, 1, 1)]:
        b.canon(idx)
    return a, b, b.copy_structure()


def abc_12(t=D):  # libxm test.c:1089-1190
    sp = [2, 7, 8]
    a = Spec((11,) * 4, {k: sp for k in range(4)}, t)
    b = Spec((11,) * 4, {k: sp for k in range(4)}, t)
    _symmetric([a, b], PERMS8, guard=b)
    c = Spec((11, 11), {0: sp, 1: sp}, t).all_canonical()
    return a, b, c


def ab_1(t=D):  # libxm test.c:418-437
    sp = {1: [1], 3: [2], 5: [3], 7: [2, 5]}
    dims = (1, 2, 3, 4, 5, 6, 7, 8)
    return (Spec(dims, sp, t).all_canonical(), Spec(dims, sp, t).all_canonical())


def ab_2(t=D):  # libxm test.c:439-456
    sp = {k: [2] for k in range(4)}
    return (Spec((4,) * 4, sp, t).all_canonical(),
            Spec((4,) * 4, sp, t).all_canonical())


def ab_3(t=D):  # libxm test.c:458-477
    sp = {0: [4], 1: [4]}
    a = Spec((8, 8), sp, t).all_canonical()
    b = Spec((8, 8), sp, t).canon((0, 0)).canon((1, 1))
    b.deriv((0, 1), (0, 0), (1, 0), -4)
    return a, b


def ab_4(t=D):  # libxm test.c:479-498
    sp = {0: [10], 1: [10]}
    a = Spec((20, 20), sp, t).canon((0, 0)).canon((1, 1))
    b = Spec((20, 20), sp, t).canon((0, 0))
    b.deriv((1, 1), (0, 0), (1, 0), -10)
    return a, b


def ab_5(t=D):  # libxm test.c:500-521
    sp = {0: [11], 1: [15]}
    a = Spec((17, 31), sp, t).canon((0, 0)).canon((1, 1))
    b = Spec((17, 31), sp, t)
    for idx in [(0, 0), (1, 1), (0, 1), (1, 0)]:
        b.canon(idx)
    return a, b


# libxm test.c:1912-1959
CONTRACT_TESTS = [
    (abc_1, "ik", "kj", "ij"), (abc_1, "ik", "kj", "ji"),
    (abc_1, "ik", "jk", "ij"), (abc_1, "ik", "jk", "ji"),
    (abc_1, "ki", "kj", "ij"), (abc_1, "ki", "kj", "ji"),
    (abc_1, "ki", "jk", "ij"), (abc_1, "ki", "jk", "ji"),
    (abc_2, "ik", "kj", "ij"), (abc_2, "ik", "kj", "ji"),
    (abc_2, "ik", "jk", "ij"), (abc_2, "ik", "jk", "ji"),
    (abc_2, "ki", "kj", "ij"), (abc_2, "ki", "kj", "ji"),
    (abc_2, "ki", "jk", "ij"), (abc_2, "ki", "jk", "ji"),
    (abc_3, "abcdefgh", "abcdefgi", "ih"),
    (abc_3, "abcdefgi", "abcdefgh", "ih"),
    (abc_3, "abcdxfgh", "abcdyfgh", "xy"),
    (abc_3, "abcdefgh", "obcdefgh", "ao"),
    (abc_3, "abcdefgh", "obcdefgh", "oa"),
    (abc_4, "abcdef", "abcijk", "ijkdef"),
    (abc_4, "abcdef", "aibjck", "ijkdef"),
    (abc_4, "badcfe", "xyzdef", "xyzabc"),
    (abc_5, "ik", "kj", "ij"), (abc_5, "ik", "kj", "ji"),
    (abc_5, "ik", "jk", "ij"), (abc_5, "ik", "jk", "ji"),
    (abc_5, "ki", "kj", "ij"), (abc_5, "ki", "kj", "ji"),
    (abc_5, "ki", "jk", "ij"), (abc_5, "ki", "jk", "ji"),
    (abc_6, "y", "xz", "xyz"),
    (abc_7, "abc", "abd", "dc"),
    (abc_8, "ab", "b", "a"),
    (abc_9, "ijab", "abcd", "ijcd"),
    (abc_9, "ijcd", "abcd", "ijab"),
    (abc_10, "abcd", "ijab", "ijcd"),
    (abc_10, "cdab", "ijab", "ijcd"),
    (abc_11, "abcd", "ijcd", "ijab"),
    (abc_12, "abcf", "abce", "ef"),
    (abc_12, "abfc", "abce", "ef"),
    (abc_12, "fabc", "bace", "ef"),
    (abc_12, "abcf", "acbe", "fe"),
    (abc_12, "afbc", "eabc", "fe"),
    (abc_12, "afcb", "bace", "fe"),
]

# libxm test.c:1878-1888
ADD_TESTS = [
    (ab_1, "abcdefgh", "abcdefgh"),
    (ab_2, "ijkl", "ijkl"), (ab_2, "ijkl", "ijlk"),
    (ab_2, "ijkl", "jkil"), (ab_2, "ijkl", "lkji"),
    (ab_3, "ij", "ij"), (ab_3, "ji", "ij"),
    (ab_4, "ij", "ij"), (ab_4, "ij", "ji"),
]

# libxm test.c:1890-1897 (mul/div) and 1899-1910 (dot)
MULDIV_TESTS = [
    (ab_1, "abcdefgh", "abcdefgh"),
    (ab_2, "ijkl", "ijkl"), (ab_2, "ijkl", "ijlk"),
    (ab_2, "ijkl", "jkil"), (ab_2, "ijkl", "lkji"),
    (ab_5, "ij", "ij"),
]
DOT_TESTS = ADD_TESTS + [(ab_5, "ij", "ij")]

# The (alpha, beta) combinations libxm runs for each case: zero or a
# random value each (libxm test.c:2021-2031).
ALPHA_BETA = [(0.0, 0.0), (0.0, "r"), ("r", 0.0), ("r", "r")]


def case_id(case):
    fn, *idx = case
    return "%s-%s" % (fn.__name__, "-".join(idx))
