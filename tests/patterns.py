"""M's alternative sparsity patterns built with scipy, independently of the
library's host builder: the reference's GetSparsityPattern flags
(cuFormatConversion.h:598-974) 0 = A^T, 1 = E + A, 2 = E + A + A^T."""
import numpy as np
import scipy.sparse as sp


def pattern_csc(n, colptr, rowidx, flag):
    """(colptr, rowidx) of the pattern, rows sorted; flag -1 = A itself."""
    B = sp.csc_matrix((np.ones(len(rowidx)), np.asarray(rowidx), np.asarray(colptr)), shape=(n, n))
    if flag == -1:
        P = B
    elif flag == 0:
        P = B.T
    elif flag == 1:
        P = B + sp.identity(n, format="csc")
    elif flag == 2:
        P = B + B.T + sp.identity(n, format="csc")
    else:
        raise ValueError(flag)
    P = sp.csc_matrix(P)
    P.sum_duplicates()
    P.sort_indices()
    return P.indptr.astype(np.int32), P.indices.astype(np.int32)
