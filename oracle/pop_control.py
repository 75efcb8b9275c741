"""numpy restatement of ipie's single-rank population control (test oracle).

Test infrastructure only: imported by tests/, never by the product package.

Restates ipie/walkers/pop_controller.py for one rank (comm.size == 1): the
weight rescaling of ``PopController.pop_control`` (:78-126), ``pair_branch``
(:304-402), ``comb`` (:199-301) and ``stochastic_reconfiguration`` (:405-450),
with the walker buffer of ``get_buffer`` / ``set_buffer`` (:128-196) as the
fields of ``BUFF_NAMES``.  The reference draws its uniforms from
``numpy.random`` inside the loops; here they are passed in (``uniforms``,
consumed in the same order: one per branched pair, one per comb, one per
reconfiguration), so that a device implementation fed the same numbers must
reproduce every decision.
"""

import numpy

# walker fields carried by a walker buffer, in ipie's order
# (ipie/walkers/base_walkers.py:128-136, uhf_walkers.py:94)
BUFF_NAMES = ["weight", "unscaled_weight", "phase", "hybrid_energy", "ovlp", "sgn_ovlp", "log_ovlp", "phia",
              "phib"]


def rescale(weight, target_weight, pairwise=True):
    """Total weight, scale and the rescaled weights of PopController.pop_control.

    Restates ipie/walkers/pop_controller.py:82-115 for one rank: total =
    sum |w| (numpy.sum, pairwise, for pair_branch and SR; the builtin sum,
    left to right, of the gathered |w| for comb: the same number up to the
    last bits); weight <- weight / scale with scale = total / target;
    ValueError when total < 1e-8.  Returns (total, unscaled, scaled)."""
    weights = numpy.abs(numpy.asarray(weight))
    total = float(numpy.sum(weights)) if pairwise else float(sum(weights))
    if total < 1e-8:
        raise ValueError(f"total weight is {total:13.8e}")
    scale = total / target_weight
    return total, numpy.asarray(weight).copy(), numpy.asarray(weight) / scale


def pair_branch_plan(weight, min_weight, max_weight, uniforms):
    """Pairing decisions of pair_branch for one rank.

    Restates ipie/walkers/pop_controller.py:304-402: walkers sorted by |w|
    (stable mergesort), the lightest and heaviest paired while the lightest is
    below min_weight or the heaviest above max_weight; with r the next uniform,
    the heavy walker is cloned over the light one if r < w_e / (w_s + w_e), else
    the light one over the heavy one (status 2 branched, 0 killed); the
    branched walker's weight becomes (w_s + w_e) / 2 before its buffer is sent.

    Which killed slot receives which branched walker follows the messages: on
    one rank every send and receive carries tag 0 (:398-420 build the tags from
    rank indices), sends are posted in slot order of the branched walkers and
    receives in slot order of the killed ones, and MPI keeps messages of one
    source and tag in order, so the k-th branched walker (ascending slot) fills
    the k-th killed slot with its buffer, weight included.  Returns (src,
    new_weight, npairs): slot i takes the buffer of walker src[i] (i itself
    when untouched) and new_weight[i] is its weight afterwards."""
    w = numpy.abs(numpy.asarray(weight, dtype=numpy.float64))
    nw = len(w)
    order = numpy.argsort(w, kind="mergesort")
    ws = w[order].copy()
    status = numpy.ones(nw, dtype=numpy.int64)
    new_weight = numpy.asarray(weight, dtype=numpy.float64).copy()
    s, e, k = 0, nw - 1, 0
    while s < e:
        if ws[s] < min_weight or ws[e] > max_weight:
            wab = ws[s] + ws[e]
            r = uniforms[k]
            k += 1
            i_s, i_e = order[s], order[e]
            if r < ws[e] / wab:
                branched, killed = i_e, i_s  # clone the heavy walker, kill the light one
            else:
                branched, killed = i_s, i_e
            status[branched] = 2
            status[killed] = 0
            new_weight[branched] = 0.5 * wab
            s += 1
            e -= 1
        else:
            break
    src = numpy.arange(nw)
    b = numpy.where(status == 2)[0]
    d = numpy.where(status == 0)[0]
    src[d] = b
    new_weight[d] = new_weight[b]
    return src, new_weight, k


def comb_plan(weight, target_weight, r):
    """Parent counts and clone/kill pairs of comb for one rank.

    Restates ipie/walkers/pop_controller.py:212-231 and 250-301: comb points
    (i + r) total / target for i < target against the cumulative weights;
    walkers with no parent point are killed, walkers with more than one are
    cloned, and the i-th clone fills the i-th killed slot (zip: each clone is
    copied once, as in the reference); every weight is then 1.  Returns (src,
    new_weight, parent_ix)."""
    w = numpy.asarray(weight, dtype=numpy.float64)
    nw = len(w)
    total = numpy.sum(w)
    cprobs = numpy.cumsum(w)
    comb = [(i + r) * (total / target_weight) for i in range(target_weight)]
    parent_ix = numpy.zeros(nw, dtype=numpy.int64)
    iw = ic = 0
    while ic < len(comb):
        if comb[ic] < cprobs[iw]:
            parent_ix[iw] += 1
            ic += 1
        else:
            iw += 1
    kill = numpy.where(parent_ix == 0)[0]
    clone = numpy.where(parent_ix > 1)[0]
    src = numpy.arange(nw)
    for c, k in zip(clone, kill):
        src[k] = c
    return src, numpy.ones(nw), parent_ix


def reconfiguration_plan(weight, zeta):
    """Walker selection of stochastic_reconfiguration for one rank.

    Restates ipie/walkers/pop_controller.py:405-450: new walker i is the old
    walker searchsorted(cumsum |w|, (i + zeta) / W total), all with weight
    total / W.  Returns (src, new_weight)."""
    cw = numpy.cumsum(numpy.abs(numpy.asarray(weight, dtype=numpy.float64)))
    nw = len(cw)
    total = cw[-1]
    src = numpy.array([numpy.searchsorted(cw, (i + zeta) / nw * total) for i in range(nw)])
    return src, numpy.full(nw, total / nw)


def apply_plan(fields, src, new_weight):
    """Copy walker buffers: slot i takes walker src[i]'s fields (BUFF_NAMES,
    from the state before any copy), then weight <- new_weight.  ``fields`` is
    a dict of arrays with the walker index first; returns a new dict."""
    out = {k: numpy.asarray(v)[src].copy() for k, v in fields.items()}
    out["weight"] = numpy.asarray(new_weight, dtype=numpy.float64).copy()
    return out


def pop_control(fields, method, target_weight, uniforms, min_weight=0.1, max_weight=4.0, reconfigure=True):
    """One single-rank PopController.pop_control call on walker ``fields``.

    Restates ipie/walkers/pop_controller.py:78-126: rescale (unscaled_weight
    keeps the weights before the division), then the method's plan and the
    buffer copies.  ``reconfigure`` says whether this call is the
    reconfiguration_freq-th one (SR).  Returns (fields, total_weight)."""
    total, unscaled, scaled = rescale(fields["weight"], target_weight, pairwise=method != "comb")
    f = dict(fields)
    f["unscaled_weight"] = unscaled
    f["weight"] = scaled
    if method == "pair_branch":
        src, nwt, _ = pair_branch_plan(scaled, min_weight, max_weight, uniforms)
        return apply_plan(f, src, nwt), total
    if method == "comb":
        src, nwt, _ = comb_plan(scaled, int(target_weight), uniforms[0])
        return apply_plan(f, src, nwt), total
    if method == "stochastic_reconfiguration":
        if not reconfigure:
            return f, total
        # the reference's buffer carries the weight first; SR overwrites it
        src, nwt = reconfiguration_plan(scaled, uniforms[0])
        return apply_plan(f, src, nwt), total
    raise ValueError(f"unknown population control method {method}")
