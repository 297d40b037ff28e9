"""The groups' GF(4) forms (CPU): the representation by 3x3 matrices over
GF(4) that the field sweep kernel multiplies instead of gathering from the
table, checked against the group files' own multiplication tables (read by
the oracle, group.c:13-81) with an independent GF(4) matrix product."""
import numpy as np
import pytest

# GF(4) = {0, 1, w, w^2}: e1 w + e0 coded 2 e1 + e0; w^2 = w + 1; sum = xor
_GM = np.zeros((4, 4), np.uint8)
for _a in range(4):
    for _b in range(4):
        # (a1 w + a0)(b1 w + b0) = a1 b1 w^2 + (a1 b0 + a0 b1) w + a0 b0
        a1, a0, b1, b0 = _a >> 1, _a & 1, _b >> 1, _b & 1
        c1 = (a1 & b1) ^ (a1 & b0) ^ (a0 & b1)
        c0 = (a1 & b1) ^ (a0 & b0)
        _GM[_a, _b] = 2 * c1 + c0


def gf4_matmul(A, B):
    """A [..., 3, 3] times B [..., 3, 3] over GF(4)."""
    A, B = np.broadcast_arrays(A, B)
    C = np.zeros(A.shape, np.uint8)
    for k in range(3):
        C ^= _GM[A[..., :, k][..., :, None], B[..., k, :][..., None, :]]
    return C


def gf4_det(M):
    e = lambda r, c: M[..., r, c]
    m = lambda a, b: _GM[a, b]
    return (m(e(0, 0), m(e(1, 1), e(2, 2)) ^ m(e(1, 2), e(2, 1))) ^
            m(e(0, 1), m(e(1, 0), e(2, 2)) ^ m(e(1, 2), e(2, 0))) ^
            m(e(0, 2), m(e(1, 0), e(2, 1)) ^ m(e(1, 1), e(2, 0))))


def test_gf4_arithmetic():
    """The field itself: w has order 3, w^2 = w + 1, every nonzero element
    has an inverse, multiplication distributes over xor."""
    w = 2
    assert _GM[w, w] == 3 and _GM[_GM[w, w], w] == 1
    for a in range(1, 4):
        assert sorted(_GM[a, 1:]) == [1, 2, 3]
    for a in range(4):
        for b in range(4):
            for c in range(4):
                assert _GM[a, b ^ c] == _GM[a, b] ^ _GM[a, c]


@pytest.mark.parametrize("name", ["mys1080-v4", "S216", "S108"])
def test_field_form_represents_the_table(dsa, oracle, name):
    """mat[x] mat[y] = mat[x y] for every pair of the file's table, det = 1,
    identity to identity, and the first two rows name each element."""
    g = dsa.load_group(dsa.group_path(name))
    mat, key = g.field()
    og = oracle.OGroup(dsa.group_path(name))
    mult, inv, _, _, _ = og.tables()
    P = g.P
    assert np.all(gf4_det(mat) == 1)
    assert np.array_equal(mat[og.id], np.eye(3, dtype=np.uint8))
    for x in range(P):
        prod = gf4_matmul(mat[x][None], mat)
        assert np.array_equal(prod, mat[mult[x]]), f"{name}: row {x} of the table"
    # inverses map to inverse matrices
    assert np.all(gf4_matmul(mat, mat[inv]) == np.eye(3, dtype=np.uint8))
    codes = (mat.reshape(P, 9)[:, :6].astype(np.uint32) << (2 * np.arange(6, dtype=np.uint32))).sum(1)
    assert len(set(codes.tolist())) == P
    assert np.array_equal(key[codes], np.arange(P))
    assert np.count_nonzero(key != 0xFFFF) == P


@pytest.mark.parametrize("name", ["myBI", "myBT"])
def test_groups_without_a_field_form(dsa, name):
    """The load-time search finds no GF(4) form for the binary icosahedral
    and tetrahedral groups; the call says so (and the field kernel is refused
    for them)."""
    g = dsa.load_group(dsa.group_path(name))
    with pytest.raises(dsa.DSAError):
        g.field()
