"""Hand-derived known answers for the BVH2 -> BVH8 collapse (CPU oracle, and
the GPU collapse on the same trees).

Each tree below is built by hand with chosen boxes, and the expected content
of every BVH8 slot is derived in the comments from two rules of the
reference:

  collapse (NexusBVH/src/Cuda/WideConverter.cu:295-330): push both children
  of the node being expanded, the child of smaller area (strict <, so on a tie
  the right child) into the slot that was popped and the other one at the end;
  then pop the inner child in the highest slot, until 8 children are held or
  none of them is inner;

  greedy slots (WideConverter.cu:25-30, 106-153): for each child in push order,
  off = (parent min + parent max) - (child min + child max) and the free slot s
  maximising (+-off.x) + (+-off.y) + (+-off.z) is taken, the sign of an axis
  being - when its bit of s is set (x: bit 2, y: bit 1, z: bit 0); strict >, so
  the lowest slot wins a tie.

Areas are half surface areas, dx*dy + dy*dz + dz*dx. Every coordinate is a
small dyadic number, so all areas and offsets are exact in FP32 and the
derivations hold bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as O

INV = 0xFFFFFFFF


def make_bvh2(leaves, tree):
    """BVH2 over leaf boxes (leaf i = primitive i, 6 floats min xyz / max xyz)
    with the shape of `tree` (nested pairs of leaf ids). Internal nodes are
    numbered in post-order from n, so the root is 2n-2 as in the library."""
    leaves = np.asarray(leaves, dtype=np.float32).reshape(-1, 6)
    n = len(leaves)
    nodes = np.zeros(2 * n - 1, dtype=O.BVH2_NODE_DTYPE)
    nodes["bmin"][:n], nodes["bmax"][:n] = leaves[:, :3], leaves[:, 3:]
    nodes["left"][:n] = INV
    nodes["right"][:n] = np.arange(n)
    nxt = [n]

    def rec(t):
        if isinstance(t, int):
            return t
        a, b = rec(t[0]), rec(t[1])
        k = nxt[0]
        nxt[0] += 1
        nodes["bmin"][k] = np.minimum(nodes["bmin"][a], nodes["bmin"][b])
        nodes["bmax"][k] = np.maximum(nodes["bmax"][a], nodes["bmax"][b])
        nodes["left"][k], nodes["right"][k] = a, b
        return k

    assert rec(tree) == 2 * n - 2
    return nodes


def prims_under(b8, prim, k):
    out = set()
    node = b8[k]
    ki = 0
    for i in range(8):
        m = int(node["meta"][i])
        if m == 0:
            continue
        if (int(node["imask"]) >> i) & 1:
            out |= prims_under(b8, prim, int(node["child_base"]) + ki)
            ki += 1
        else:
            out.add(int(prim[int(node["prim_base"]) + (m & 31)]))
    return out


def slots(b8, prim, k=0):
    """{slot: frozenset of the primitives below that slot} of BVH8 node k,
    and {slot: index of the inner child's BVH8 node}."""
    node = b8[k]
    got, child = {}, {}
    ki = 0
    for i in range(8):
        m = int(node["meta"][i])
        if m == 0:
            continue
        if (int(node["imask"]) >> i) & 1:
            c = int(node["child_base"]) + ki
            ki += 1
            got[i] = frozenset(prims_under(b8, prim, c))
            child[i] = c
        else:
            got[i] = frozenset([int(prim[int(node["prim_base"]) + (m & 31)])])
    return got, child


def fs(*ids):
    return frozenset(ids)


def slab(x0, x1):
    """Box [x0, x1] x [0, 1] x [0, 1]: every such child has off.y = off.z = 0
    inside a parent with the same y/z extent, so only the x sign decides
    (bit 2 clear for children left of the parent's centre)."""
    return [x0, 0, 0, x1, 1, 1]


# --------------------------------------------------------------------- trees
# (1) Highest-inner-slot pop. R = (S, T); S has 4 leaves, T 8.
#   leaves (x extents, y/z [0,1]): i=0 [0,1] j=1 [.5,2] k=2 [2,3] l=3 [3,4.5]
#   a=4 [10,11] b=5 [11,12] c=6 [12,13] d=7 [13,15]
#   e=8 [20,21] f=9 [21,23] g=10 [23,24] h=11 [24,27]
#   S1=(i,j) [0,2]  S2=(k,l) [2,4.5]  S=[0,4.5]
#   T11=(a,b) [10,12]  T12=(c,d) [12,15]  T1=[10,15]
#   T21=(e,f) [20,23]  T22=(g,h) [23,27]  T2=[20,27]  T=[10,27]
# With dy = dz = 1 the area is 2*dx + 1, so area order is x-extent order.
# Root collapse (slot: node; inner slots starred):
#   push S, T: S smaller -> 0:S* 1:T*                               count 2
#   pop 1 (T):  T1 < T2   -> 1:T1* 2:T2*                            count 3
#   pop 2 (T2): T21 < T22 -> 2:T21* 3:T22*                          count 4
#   pop 3 (T22): g < h    -> 3:g 4:h                                count 5
#   pop 2 (T21): e < f    -> 2:e 5:f                                count 6
#   pop 1 (T1): T11 < T12 -> 1:T11* 6:T12*                          count 7
#   pop 6 (T12): c < d    -> 6:c 7:d                                count 8, stop
# Push order: S, T11, e, g, h, f, c, d.
# Greedy in R = [0,27]: pc.x = 27; off.x = 27 - (x0 + x1):
#   S 22.5 -> 0;  T11 5 -> 1;  e -14 -> 4;  g -20 -> 5;  h -24 -> 6;  f -17 -> 7;
#   c 2 -> 2 (0, 1 taken);  d -1 -> 3 (4..7 taken; 3 is the only free slot).
# Child S (pc.x 4.5): S1 (area 5) < S2 (6) -> 0:S1* 1:S2*; pop 1: k < l ->
#   1:k 2:l; pop 0: i < j -> 0:i 3:j. Push order i, k, l, j; off.x:
#   i 3.5 -> 0; k -0.5 -> 4; l -3 -> 5; j 2 -> 1.
# Child T11 (pc.x 22): a, b tie -> right child b first: push order b, a;
#   b: off -1 -> 4; a: off 1 -> 0.
# Popping the LOWEST inner slot instead would give the root
#   {i}, {a}, {k}, {j}, T2, T12, {b}, {l}.
POP_LEAVES = [slab(0, 1), slab(.5, 2), slab(2, 3), slab(3, 4.5), slab(10, 11), slab(11, 12), slab(12, 13),
              slab(13, 15), slab(20, 21), slab(21, 23), slab(23, 24), slab(24, 27)]
POP_TREE = (((0, 1), (2, 3)), (((4, 5), (6, 7)), ((8, 9), (10, 11))))
POP_ROOT = {0: fs(0, 1, 2, 3), 1: fs(4, 5), 2: fs(6), 3: fs(7), 4: fs(8), 5: fs(10), 6: fs(11), 7: fs(9)}
POP_S = {0: fs(0), 1: fs(1), 4: fs(2), 5: fs(3)}
POP_T11 = {0: fs(4), 4: fs(5)}

# (2) Area tie at the root: R = (A, B), two 7-leaf caterpillars of equal
# area (B is A moved by +10 in x). A = (a1, (a2, (a3, (a4, (a5, (a6, a7)))))),
# a_t = [t-1, t] (ids 0..6), B likewise with ids 7..13 on [10, 17]; each
# leaf is smaller than the inner remainder beside it, and (a6, a7) tie.
#   push A, B: tie -> right child B into the popped slot: 0:B* 1:A*
#   pop 1: 1:a1 2:A'*; pop 2: 2:a2 3:A''*; pop 3: 3:a3 4:*; pop 4: 4:a4 5:*;
#   pop 5: 5:a5 6:(a6,a7)*; pop 6: tie -> 6:a7 7:a6; count 8.
# Push order B, a1, a2, a3, a4, a5, a7, a6. Greedy, pc.x = 17:
#   B -10 -> 4; a1 16 -> 0; a2 14 -> 1; a3 12 -> 2; a4 10 -> 3;
#   a5 8 -> 5 (0..3 taken, the free slots all cost -8, lowest wins);
#   a7 4 -> 6; a6 6 -> 7.
# A tie sent to the left child would expand B instead: {A, b1..b7}.
TIE_LEAVES = [slab(t - 1, t) for t in range(1, 8)] + [slab(t + 9, t + 10) for t in range(1, 8)]
TIE_TREE = (
    (0, (1, (2, (3, (4, (5, 6)))))),
    (7, (8, (9, (10, (11, (12, 13)))))),
)
TIE_ROOT = {0: fs(0), 1: fs(1), 2: fs(2), 3: fs(3), 4: fs(*range(7, 14)), 5: fs(4), 6: fs(6), 7: fs(5)}

# (3) Greedy ties and push order: L0 = L1 = [0,1]^3, L2 = [3,4] x [0,1]^2,
# R = ((L0, L1), L2). Areas: (L0, L1) 3 = L2 3, tie -> 0:L2 1:(L0,L1)*;
# pop 1: L0, L1 tie -> 1:L1 2:L0. Push order L2, L1, L0. pc.x = 4:
#   L2 off -3 -> 4 (slots 4..7 tie at 3, lowest); L1 off 3 -> 0; L0 -> 1.
GREEDY_TIE_LEAVES = [slab(0, 1), slab(0, 1), slab(3, 4)]
GREEDY_TIE_TREE = ((0, 1), 2)
GREEDY_TIE_ROOT = {0: fs(1), 1: fs(0), 4: fs(2)}

# (4) A child whose best slot is taken: parent [0,4]^3, C0 = [0,1]^3,
# C1 = [3,4]^3, C2 = [2,3] x [3,4] x [3,4]; R = (C0, (C1, C2)).
# Areas: C0 3 < (C1, C2) = [2,4]x[3,4]^2: 2+1+2 = 5 -> 0:C0 1:M*;
# pop 1: C1 3 = C2 3 tie -> 1:C2 2:C1. Push order C0, C2, C1. pc = (4,4,4):
#   C0 off (3,3,3)    -> 0 (cost 9);
#   C2 off (-1,-3,-3) -> 7 (cost 7);
#   C1 off (-3,-3,-3) -> 7 is taken; slots 3, 5, 6 cost 3 -> 3 (lowest).
BLOCKED_LEAVES = [[0, 0, 0, 1, 1, 1], [3, 3, 3, 4, 4, 4], [2, 3, 3, 3, 4, 4]]
BLOCKED_TREE = (0, (1, 2))
BLOCKED_ROOT = {0: fs(0), 3: fs(1), 7: fs(2)}

HAND_CASES = {
    "pop_highest": (POP_LEAVES, POP_TREE, POP_ROOT),
    "area_tie": (TIE_LEAVES, TIE_TREE, TIE_ROOT),
    "greedy_tie": (GREEDY_TIE_LEAVES, GREEDY_TIE_TREE, GREEDY_TIE_ROOT),
    "blocked_slot": (BLOCKED_LEAVES, BLOCKED_TREE, BLOCKED_ROOT),
}


def octant_case(seed):
    """Eight leaves, one per octant of the parent [-1,1]^3 (the (-,-,-) leaf
    touches -1 and the (+,+,+) leaf +1, so the parent's centre is 0), under a
    random tree shape. Closed form: every leaf lands in the slot whose bits
    are the signs of its centre (x: 4, y: 2, z: 1), since that slot is each
    leaf's unique best and they are all different; the push order does not
    matter."""
    rng = np.random.default_rng(seed)
    octs = rng.permutation(8)
    boxes = []
    for o in octs:
        sgn = np.array([(o >> 2) & 1, (o >> 1) & 1, o & 1]) * 2 - 1
        c = sgn * 0.5 + rng.integers(-8, 9, 3) / 256.0
        h = rng.integers(4, 40, 3) / 256.0
        lo, hi = c - h, c + h
        if o == 0:
            lo = np.array([-1.0, -1.0, -1.0])
        if o == 7:
            hi = np.array([1.0, 1.0, 1.0])
        boxes.append(np.concatenate([lo, hi]))

    def shape(ids):
        if len(ids) == 1:
            return int(ids[0])
        cut = int(rng.integers(1, len(ids)))
        return (shape(ids[:cut]), shape(ids[cut:]))
    want = {int(o): fs(i) for i, o in enumerate(octs)}
    return boxes, shape(list(range(8))), want


def check_root(b8, prim, want):
    got, child = slots(b8, prim, 0)
    assert got == want, (got, want)
    return child


# ------------------------------------------------------------------ CPU pins
@pytest.mark.parametrize("case", sorted(HAND_CASES))
def test_oracle_collapse_hand_derived(case):
    leaves, tree, want = HAND_CASES[case]
    nodes = make_bvh2(leaves, tree)
    b8, prim = O.collapse_bvh8(nodes, len(leaves))
    child = check_root(b8, prim, want)
    if case == "pop_highest":
        assert slots(b8, prim, child[0])[0] == POP_S
        assert slots(b8, prim, child[1])[0] == POP_T11


@pytest.mark.parametrize("seed", range(12))
def test_oracle_greedy_octant_closed_form(seed):
    boxes, tree, want = octant_case(seed)
    b8, prim = O.collapse_bvh8(make_bvh2(boxes, tree), 8)
    assert len(b8) == 1
    check_root(b8, prim, want)


def test_oracle_greedy_two_children_sides():
    """Two leaves: the one left of the parent's centre in x takes slot 0 and
    the right one slot 4 (ties over y/z go to the lowest slot); in y, slots 0
    and 2; in z, slots 0 and 1; a child at the centre in every axis (off = 0,
    all eight slots tie) takes slot 0."""
    for axis, right_slot in ((0, 4), (1, 2), (2, 1)):
        lo = np.zeros(6)
        lo[3:] = 1
        hi = lo.copy()
        hi[axis] += 2
        hi[axis + 3] += 2
        for tree in ((0, 1), (1, 0)):
            b8, prim = O.collapse_bvh8(make_bvh2([lo, hi], tree), 2)
            check_root(b8, prim, {0: fs(0), right_slot: fs(1)})
    # centred child: [0,4]^3 around [1,3]^3 (the big box is pushed second and
    # sees only slots 1..7 free, where it also has off = 0 -> slot 1)
    b8, prim = O.collapse_bvh8(make_bvh2([[1, 1, 1, 3, 3, 3], [0, 0, 0, 4, 4, 4]], (0, 1)), 2)
    check_root(b8, prim, {0: fs(0), 1: fs(1)})


# ------------------------------------------------------------------ GPU pins
@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(HAND_CASES))
def test_gpu_collapse_hand_derived(case):
    from nexusbvh_amd import stages
    from nexusbvh_amd.native import DeviceBuffer, check
    leaves, tree, want = HAND_CASES[case]
    nodes = make_bvh2(leaves, tree)
    n = len(leaves)
    d = DeviceBuffer.from_array(nodes.view(np.uint8))
    st, b8d, primd, count = stages.collapse(d, n)
    check(st, "collapse")
    b8 = stages.bvh8_to_host(b8d, count)
    prim = primd.to_array(np.uint32, n)
    child = check_root(b8, prim, want)
    if case == "pop_highest":
        assert slots(b8, prim, child[0])[0] == POP_S
        assert slots(b8, prim, child[1])[0] == POP_T11


@pytest.mark.gpu
def test_gpu_greedy_octant_closed_form():
    from nexusbvh_amd import stages
    from nexusbvh_amd.native import DeviceBuffer, check
    for seed in range(12):
        boxes, tree, want = octant_case(seed)
        d = DeviceBuffer.from_array(make_bvh2(boxes, tree).view(np.uint8))
        st, b8d, primd, count = stages.collapse(d, 8)
        check(st, "collapse")
        assert count == 1
        check_root(stages.bvh8_to_host(b8d, count), primd.to_array(np.uint32, 8), want)


def test_quantized_face_can_round_inward_past_the_child():
    """The hazard the traversal's box widening must cover (CPU, oracle BVH8):
    with p.x = -1000 and a grid step of 8, a child whose min x is -2^-20 gets
    qlo = floor(fl(1000 - 2^-20) / 8) = 125, so its decoded face
    -1000 + 125 * 8 = 0 lies right of the child. A widening of
    (|l| + |p|) * 2^-21 reaches past it; one of |l| * 2^-21 alone (0 here)
    does not."""
    e = np.float32(2.0 ** -20)
    tris = np.array([[-e, 1, -1, -e, 1, 1, 5, 1, 0],
                     [-1000, 50, 50, -999, 50, 50, -1000, 51, 50],
                     [1000, -50, -50, 999, -50, -50, 1000, -51, -50]], dtype=np.float32)
    ref = O.build(tris, 1, 32, 32, 8)
    b8, prim = ref["bvh8"], ref["prim_idx"]
    assert len(b8) == 1
    node = b8[0]
    assert node["p"][0] == np.float32(-1000) and int(node["e"][0]) == 130  # step 2^(130-127) = 8
    slot = [i for i in range(8) if node["meta"][i] and prim[int(node["prim_base"]) + (node["meta"][i] & 31)] == 0]
    assert len(slot) == 1
    q = int(node["qlox"][slot[0]])
    assert q == 125
    lx = np.float32(np.float32(-1000) + np.float32(q) * np.float32(8))
    assert lx == 0 and lx > -e
    assert lx - (abs(lx) + np.float32(1000)) * np.float32(2.0 ** -21) < -e
    assert lx - abs(lx) * np.float32(2.0 ** -21) > -e
