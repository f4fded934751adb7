"""Host-side data preparation of the reference, restated for the tests:
row exclusion (exclusion list, covariate NaN, interaction NaN;
src/AIOwrapper.cpp:229-253,434-470) and interaction synthesis
(src/AIOwrapper.cpp:767-840).  Produces the matrices the FP64 oracle scans
and the descriptors the GPU library takes."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_fixture():
    z = np.load(os.path.join(GOLDEN, "interactions_subset.npz"))
    return {k: z[k] for k in z.files}


def exclusion_keep(n, text):
    from oracle import oracle as O
    rc, keep = O.exclusions(text, n)
    assert rc >= 0, rc
    return keep


def prepare(fx, inter_name=None, multi=False, keep_dep=False, excl_text=None, limit_t=50, limit_m=100,
            dosage=None):
    """Returns dict(XL, XR (n x r*m, NaN kept), Y, r, m, E, col_src, col_env, G, dos).

    dosage: (factors, file_r, summed) -- the genotype file holds file_r
    columns per SNP and the model's columns (oracle.dosage_apply) take the
    place of the file columns; G then holds the raw file columns the GPU
    library combines itself."""
    XL = fx["XL"].astype(np.float64)
    G = fx["XR"][:, :limit_m].astype(np.float64)
    Y = fx["Y"][:, :limit_t].astype(np.float64)
    n = XL.shape[0]
    keep = np.ones(n, bool) if excl_text is None else exclusion_keep(n, excl_text)
    keep &= ~np.isnan(XL).any(axis=1)
    E = None
    if inter_name is not None:
        E = fx["E_" + inter_name].astype(np.float64)
        keep &= ~np.isnan(E).any(axis=1)
        E = E[keep]
    XL, G, Y = XL[keep], G[keep], Y[keep]
    G_raw = G
    if dosage is not None:
        from oracle import oracle as O
        factors, file_r, summed = dosage
        G = G[:, :(G.shape[1] // file_r) * file_r]
        G_raw = G
        G = O.dosage_apply(G, file_r, factors, summed)
        if not summed and len(factors) > 1:
            assert E is None, "the reference refuses interactions with more than one column per SNP"
    m_file = G.shape[1]
    if E is None:
        r = 1 if dosage is None or dosage[2] else len(dosage[0])
        m = m_file // r
        col_src = np.arange(m * r, dtype=np.int32)
        col_env = -np.ones(m * r, dtype=np.int32)
    elif not multi:
        k = E.shape[1]
        r, m = k + int(keep_dep), m_file
        src, env = [], []
        for s in range(m_file):
            for h in range(k):
                src.append(s); env.append(h)
            if keep_dep:
                src.append(s); env.append(-1)
        col_src, col_env = np.array(src, np.int32), np.array(env, np.int32)
    else:
        k = E.shape[1]
        r = 1 + int(keep_dep)
        fb = max(1, min(m_file, 500) // k)
        src, env = [], []
        for b0 in range(0, m_file, fb):
            bs = min(fb, m_file - b0)
            for h in range(k):
                for i in range(bs):
                    src.append(b0 + i); env.append(h)
                    if keep_dep:
                        src.append(b0 + i); env.append(-1)
        col_src, col_env = np.array(src, np.int32), np.array(env, np.int32)
        m = m_file * k
    cols = []
    for j in range(len(col_src)):
        v = G[:, col_src[j]].copy()
        if col_env[j] >= 0:
            v = v * E[:, col_env[j]]
        cols.append(v)
    XR = np.stack(cols, axis=1)
    return dict(XL=XL, XR=XR, Y=Y, r=r, m=m, E=E, G=G_raw, col_src=col_src, col_env=col_env,
                n=int(keep.sum()), realN=n, dos=dosage)


# The reference check program's interaction scenarios (tests/test.cpp:201-308),
# no dosage model: (multi, keep, file, expected lower bound of total_sig_results).
def scenarios(limit_t=50):
    names = ["0", "1", "R", "R1", "R0", "10", "11", "011", "1RR"]
    single = [0, limit_t, 0, limit_t, 0, 0, 0, 0, limit_t]
    mult = [0, limit_t, 0, limit_t, 0, limit_t, 2 * limit_t, 2 * limit_t, limit_t]
    keep_names = ["R", "R0", "RR"]
    single_keep = [limit_t, 0, limit_t]
    mult_keep = [limit_t, limit_t, 2 * limit_t]
    out = []
    for j, multi in enumerate([False, True]):
        for nm, e1, e2 in zip(names, single, mult):
            out.append((multi, False, "INT" + nm, e2 if multi else e1))
        for nm, e1, e2 in zip(keep_names, single_keep, mult_keep):
            out.append((multi, True, "INT" + nm, e2 if multi else e1))
    return out


# Settings of that check program (tests/test.cpp:155-166), after the
# threshold normalisation of src/Algorithm.cpp:338-342.
CHECK_SETTINGS = dict(min_p_store=0.1, min_p_disp=0.00005, min_r2_store=0.000001, disp_cov=True)
