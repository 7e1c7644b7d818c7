"""Restatement of GEMMPermutationOptimizer::compute_index_list_contraction
(expressions/permopt.h:254-381) and to_cblas_op (permopt.h:47-57).

TEST INFRASTRUCTURE ONLY (part of the oracle).  Note that the constructor that
takes target indices ignores them (permopt.h:203-213), so the decision depends
on the argument index lists only.
"""
IDENTITY, MATRIX_TRANSPOSE, GENERAL = 1, 2, 3


def compute_index_list_contraction(left, right, prefer_to_permute_left=True):
    left = list(left)
    right = list(right)
    lrank, rrank = len(left), len(right)
    res_left, res_right, res = [], [], []
    for v in left:  # permopt.h:277-288
        if v not in right:
            res_left.append(v)
            res.append(v)
        else:
            res_right.append(v)
    inner_rank = len(res_right)
    left_outer_rank = len(res_left)
    right_outer_rank = rrank - inner_rank
    if inner_rank == 0:  # permopt.h:300-311
        for v in right:
            res_right.append(v)
            res.append(v)
        return res_left, res_right, res, GENERAL, GENERAL
    ordered = True
    l_no, l_tr, r_no, r_tr = True, True, True, True
    perm_left = (lrank < rrank) or (lrank == rrank and prefer_to_permute_left)
    for i, idx in enumerate(right):  # permopt.h:330-367
        if idx not in left:
            res_right.append(idx)
            res.append(idx)
            continue
        j = left.index(idx)
        x = len(res_left) - left_outer_rank
        ordered = ordered and (res_right[x] == idx)
        l_no = l_no and (j >= left_outer_rank)
        l_tr = l_tr and (j < inner_rank)
        r_no = r_no and (i < inner_rank)
        r_tr = r_tr and (i >= right_outer_rank)
        if ordered:
            res_left.append(idx)
        elif perm_left:
            res_left.append(idx)
            res_right[x] = idx
            l_no = l_tr = False
        else:
            res_left.append(res_right[x])
            r_no = r_tr = False

    def op(no, tr):
        return IDENTITY if no else (MATRIX_TRANSPOSE if tr else GENERAL)

    return res_left, res_right, res, op(l_no, l_tr), op(r_no, r_tr)


def to_cblas_op(permtype):
    """1 = Transpose for matrix_transpose, else 0 = NoTranspose."""
    return 1 if permtype == MATRIX_TRANSPOSE else 0
