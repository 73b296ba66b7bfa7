// Reverse-mode derivatives of the soft step, written by hand stage by stage
// (the reverse of m3_serial.h's order): sleep, store, then per substep the
// relax pass, integrate-positions, the solve pass, warm start and
// integrate-velocities, then prepare and the narrowphase geometry of
// plane-hull and hull-hull manifolds.  Each function takes the adjoints
// ("bars") of a stage's outputs and adds the adjoints of its inputs,
// re-evaluating the stage's forward values from the step's tape.  Every
// branch (clamps, clipping cases, sleep, which features touch) is taken as
// the forward took it: the result is the derivative of the step on the
// piece of its domain the forward point lies in, which is what finite
// differences see at contact-stable states.
//
// Scope: dynamic boxes (hulls) on static planes, no joints, restitution and
// rolling resistance zero, forces and no torques.  Continuous pull-backs
// against planes are differentiated through their time of impact; others
// are passed through as the identity and counted.
#ifndef M3_ADJOINT_H
#define M3_ADJOINT_H

#include "m3_serial.h"

// ---- adjoints of the math helpers ------------------------------------------

M3_HD m3Vec3 m3Z3(void) { return m3V3(0.0f, 0.0f, 0.0f); }
M3_HD m3Quat m3ZQ(void) { return m3Q(0.0f, 0.0f, 0.0f, 0.0f); }
M3_HD m3Quat m3AddQ(m3Quat a, m3Quat b) { return m3Q(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
M3_HD void m3Acc3(m3Vec3* a, m3Vec3 b) { *a = m3Add3(*a, b); }
M3_HD void m3AccQ(m3Quat* a, m3Quat b) { *a = m3AddQ(*a, b); }

// Transposed product M^T v.
M3_HD m3Vec3 m3MulMTV3(m3Mat3 m, m3Vec3 v)
{
    return m3V3(m3Dot3(m.cx, v), m3Dot3(m.cy, v), m3Dot3(m.cz, v));
}

// M += a b^T (column-major: column j gets a * b_j).
M3_HD void m3AccOuter(m3Mat3* m, m3Vec3 a, m3Vec3 b)
{
    m->cx = m3Add3(m->cx, m3MulSV3(b.x, a));
    m->cy = m3Add3(m->cy, m3MulSV3(b.y, a));
    m->cz = m3Add3(m->cz, m3MulSV3(b.z, a));
}

M3_HD m3real m3FrobDot(m3Mat3 a, m3Mat3 b) { return m3Dot3(a.cx, b.cx) + m3Dot3(a.cy, b.cy) + m3Dot3(a.cz, b.cz); }

// y = R(q) v: adds q̄ and returns v̄ (= R(q)^T ȳ up to the operation order).
M3_HD m3Vec3 m3AdjRotate(m3Quat q, m3Vec3 v, m3Vec3 yb, m3Quat* qb)
{
    m3Vec3 u = m3V3(q.x, q.y, q.z);
    m3Vec3 t = m3MulSV3(2.0f, m3Cross3(u, v));
    // y = v + w t + u x t
    m3Vec3 vb = yb;
    m3real wb = m3Dot3(yb, t);
    m3Vec3 tb = m3Add3(m3MulSV3(q.w, yb), m3Cross3(yb, u));
    m3Vec3 ub = m3Cross3(t, yb);
    // t = 2 u x v
    ub = m3Add3(ub, m3MulSV3(2.0f, m3Cross3(v, tb)));
    vb = m3Add3(vb, m3MulSV3(2.0f, m3Cross3(tb, u)));
    qb->x += ub.x;
    qb->y += ub.y;
    qb->z += ub.z;
    qb->w += wb;
    return vb;
}

// y = R(q)^T v = R(conj q) v.
M3_HD m3Vec3 m3AdjInvRotate(m3Quat q, m3Vec3 v, m3Vec3 yb, m3Quat* qb)
{
    m3Quat cb = m3ZQ();
    m3Vec3 vb = m3AdjRotate(m3Conj(q), v, yb, &cb);
    qb->x -= cb.x;
    qb->y -= cb.y;
    qb->z -= cb.z;
    qb->w += cb.w;
    return vb;
}

// y = a b (quaternion product): adds ā and b̄.
M3_HD void m3AdjMulQuat(m3Quat a, m3Quat b, m3Quat yb, m3Quat* ab, m3Quat* bb)
{
    // y.w = aw bw - ax bx - ay by - az bz
    // y.x = aw bx + ax bw + ay bz - az by
    // y.y = aw by - ax bz + ay bw + az bx
    // y.z = aw bz + ax by - ay bx + az bw
    ab->w += yb.w * b.w + yb.x * b.x + yb.y * b.y + yb.z * b.z;
    ab->x += -yb.w * b.x + yb.x * b.w - yb.y * b.z + yb.z * b.y;
    ab->y += -yb.w * b.y + yb.x * b.z + yb.y * b.w - yb.z * b.x;
    ab->z += -yb.w * b.z - yb.x * b.y + yb.y * b.x + yb.z * b.w;
    bb->w += yb.w * a.w + yb.x * a.x + yb.y * a.y + yb.z * a.z;
    bb->x += -yb.w * a.x + yb.x * a.w + yb.y * a.z - yb.z * a.y;
    bb->y += -yb.w * a.y - yb.x * a.z + yb.y * a.w + yb.z * a.x;
    bb->z += -yb.w * a.z + yb.x * a.y - yb.y * a.x + yb.z * a.w;
}

// y = x / |x| (above the degenerate threshold): returns x̄.
M3_HD m3Vec3 m3AdjNormalize3(m3Vec3 x, m3Vec3 yb)
{
    m3real len = m3Length3(x);
    if (len < 1.19209290e-7f)
    {
        return m3Z3();
    }
    m3real inv = 1.0f / len;
    m3Vec3 y = m3MulSV3(inv, x);
    return m3MulSV3(inv, m3Sub3(yb, m3MulSV3(m3Dot3(y, yb), y)));
}

M3_HD m3Quat m3AdjNormalizeQuat(m3Quat x, m3Quat yb)
{
    m3real mag = m3Sqrt(x.x * x.x + x.y * x.y + x.z * x.z + x.w * x.w);
    if (mag < 1.19209290e-7f)
    {
        return m3ZQ();
    }
    m3real inv = 1.0f / mag;
    m3Quat y = m3Q(inv * x.x, inv * x.y, inv * x.z, inv * x.w);
    m3real d = y.x * yb.x + y.y * yb.y + y.z * yb.z + y.w * yb.w;
    return m3Q(inv * (yb.x - d * y.x), inv * (yb.y - d * y.y), inv * (yb.z - d * y.z),
               inv * (yb.w - d * y.w));
}

// y = IntegrateRotation(q, dr): adds q̄, returns d̄r.
M3_HD m3Vec3 m3AdjIntegrateRotation(m3Quat q, m3Vec3 dr, m3Quat yb, m3Quat* qb)
{
    m3real ax = 0.5f * dr.x;
    m3real ay = 0.5f * dr.y;
    m3real az = 0.5f * dr.z;
    m3Quat qd;
    qd.w = -(ax * q.x + ay * q.y + az * q.z);
    qd.x = ax * q.w + ay * q.z - az * q.y;
    qd.y = -ax * q.z + ay * q.w + az * q.x;
    qd.z = ax * q.y - ay * q.x + az * q.w;
    m3Quat sum = m3Q(q.x + qd.x, q.y + qd.y, q.z + qd.z, q.w + qd.w);
    m3Quat sb = m3AdjNormalizeQuat(sum, yb);
    qb->x += sb.x + sb.y * az - sb.z * ay - sb.w * ax;
    qb->y += -sb.x * az + sb.z * ax - sb.w * ay;
    qb->z += sb.x * ay - sb.y * ax - sb.w * az;
    qb->w += sb.w + sb.x * ax + sb.y * ay + sb.z * az;
    qb->y += sb.y;
    qb->z += sb.z;
    m3real axb = sb.x * q.w - sb.y * q.z + sb.z * q.y - sb.w * q.x;
    m3real ayb = sb.x * q.z + sb.y * q.w - sb.z * q.x - sb.w * q.y;
    m3real azb = -sb.x * q.y + sb.y * q.x + sb.z * q.w - sb.w * q.z;
    return m3V3(0.5f * axb, 0.5f * ayb, 0.5f * azb);
}

// Impulse P applied at arms a (on A, subtracted) and b (on B, added), as
// m3ApplyImpulse.  Given the velocity bars (unchanged by the additive
// update), adds the bars of P, the arms and the mass terms; returns P̄.
typedef struct m3ConBar
{
    // adjoints of a prepared constraint's fields
    m3Vec3 normal, t1, t2, originA, originB;
    m3Vec3 rA[M3_MANIFOLD_MAX_POINTS], rB[M3_MANIFOLD_MAX_POINTS];
    m3real baseSeparation[M3_MANIFOLD_MAX_POINTS];
    m3real normalMass[M3_MANIFOLD_MAX_POINTS];
    m3real leverArm[M3_MANIFOLD_MAX_POINTS];
    m3real invMassA, invMassB;
    m3Mat3 invIA, invIB;
    m3real friction;
    m3real k11, k12, k22;
    m3real twistMass;
    // running adjoints of the accumulators
    m3real lambda[M3_MANIFOLD_MAX_POINTS];
    m3real f1, f2, twist;
} m3ConBar;

// Adjoints of one body's state.
typedef struct m3BodyBar
{
    m3Vec3 p; // position (as a real; positions themselves are f64)
    m3Quat q;
    m3Vec3 v, w;
    m3Vec3 dp; // accumulated substep deltas
    m3Quat dq;
} m3BodyBar;

M3_HD m3Vec3 m3AdjApply(const m3Constraint* c, m3Vec3 P, m3Vec3 a, m3Vec3 b, const m3BodyBar* A,
                        const m3BodyBar* B, m3ConBar* cb, m3Vec3* ab, m3Vec3* bb)
{
    m3Vec3 Pb = m3Z3();
    if (c->dynA)
    {
        // vA -= invMA P; wA -= invIA (a x P)
        Pb = m3Sub3(Pb, m3MulSV3(c->invMassA, A->v));
        cb->invMassA -= m3Dot3(P, A->v);
        m3Vec3 u = m3MulMTV3(c->invIA, A->w);
        Pb = m3Sub3(Pb, m3Cross3(u, a));
        m3Acc3(ab, m3Neg3(m3Cross3(P, u)));
        m3AccOuter(&cb->invIA, m3Neg3(A->w), m3Cross3(a, P));
    }
    if (c->dynB)
    {
        Pb = m3Add3(Pb, m3MulSV3(c->invMassB, B->v));
        cb->invMassB += m3Dot3(P, B->v);
        m3Vec3 u = m3MulMTV3(c->invIB, B->w);
        Pb = m3Add3(Pb, m3Cross3(u, b));
        m3Acc3(bb, m3Cross3(P, u));
        m3AccOuter(&cb->invIB, B->w, m3Cross3(b, P));
    }
    return Pb;
}

// ---- contact solve (m3SolveContact) -----------------------------------------

// Reverse of one m3SolveContact call.  `rec` holds the velocities and
// accumulators it started from; (dpA, dqA, dpB, dqB) the body deltas it
// read.  On entry A, B and cb's running accumulator bars hold the adjoints of
// the call's outputs; on return, of its inputs.  Constant fields' bars add
// into cb, delta bars into dA/dB.
M3_HD void m3AdjSolveContact(const m3Constraint* c, const m3TapePass* rec, m3Vec3 dpA, m3Quat dqA,
                             m3Vec3 dpB, m3Quat dqB, m3real invH, m3real pushMax, int useBias,
                             m3BodyBar* A, m3BodyBar* B, m3ConBar* cb)
{
    const int32_t np = c->pointCount;
    // Forward re-evaluation with the per-point states kept.
    m3Vec3 vA[M3_MANIFOLD_MAX_POINTS + 1], wA[M3_MANIFOLD_MAX_POINTS + 1];
    m3Vec3 vB[M3_MANIFOLD_MAX_POINTS + 1], wB[M3_MANIFOLD_MAX_POINTS + 1];
    m3real sK[M3_MANIFOLD_MAX_POINTS], biasK[M3_MANIFOLD_MAX_POINTS], vnK[M3_MANIFOLD_MAX_POINTS];
    m3real msK[M3_MANIFOLD_MAX_POINTS], isK[M3_MANIFOLD_MAX_POINTS];
    int branchK[M3_MANIFOLD_MAX_POINTS]; // 0: s > 0, 1: soft and above the push cap, 2: capped, 3: relax
    int activeK[M3_MANIFOLD_MAX_POINTS];
    m3real deltaK[M3_MANIFOLD_MAX_POINTS], newK[M3_MANIFOLD_MAX_POINTS];
    m3Vec3 dsK[M3_MANIFOLD_MAX_POINTS];
    vA[0] = rec->vA;
    wA[0] = rec->wA;
    vB[0] = rec->vB;
    wB[0] = rec->wB;
    m3real passNormal = 0.0f;
    m3real twistLimit = 0.0f;
#pragma unroll
    for (int32_t k = 0; k < M3_MANIFOLD_MAX_POINTS; ++k)
    {
        if (k >= np)
        {
            vA[k + 1] = vA[k];
            wA[k + 1] = wA[k];
            vB[k + 1] = vB[k];
            wB[k + 1] = wB[k];
            continue;
        }
        const m3CPoint* cp = &c->points[k];
        m3Vec3 rsA = m3RotateVec3(dqA, cp->rA);
        m3Vec3 rsB = m3RotateVec3(dqB, cp->rB);
        m3Vec3 ds = m3Add3(m3Sub3(dpB, dpA), m3Sub3(rsB, rsA));
        m3real s = cp->baseSeparation + m3Dot3(ds, c->normal);
        m3real bias = 0.0f;
        m3real ms = 1.0f;
        m3real is = 0.0f;
        int br = 3;
        if (s > 0.0f)
        {
            bias = s * invH;
            br = 0;
        }
        else if (useBias)
        {
            m3real soft = c->softness.biasRate * s;
            br = soft > -pushMax ? 1 : 2;
            bias = m3MaxF(soft, -pushMax);
            ms = c->softness.massScale;
            is = c->softness.impulseScale;
        }
        m3real vn = m3Dot3(m3Sub3(m3Add3(vB[k], m3Cross3(wB[k], cp->rB)), m3Add3(vA[k], m3Cross3(wA[k], cp->rA))),
                           c->normal);
        m3real lam = rec->lambda[k];
        m3real imp = -cp->normalMass * ms * (vn + bias) - is * lam;
        m3real x = lam + imp;
        m3real nl = m3MaxF(x, 0.0f);
        m3real d = nl - lam;
        sK[k] = s;
        biasK[k] = bias;
        vnK[k] = vn;
        msK[k] = ms;
        isK[k] = is;
        branchK[k] = br;
        activeK[k] = x > 0.0f;
        deltaK[k] = d;
        newK[k] = nl;
        dsK[k] = ds;
        passNormal += nl;
        twistLimit += cp->leverArm * nl;
        m3Vec3 P = m3MulSV3(d, c->normal);
        vA[k + 1] = vA[k];
        wA[k + 1] = wA[k];
        vB[k + 1] = vB[k];
        wB[k + 1] = wB[k];
        if (c->dynA)
        {
            vA[k + 1] = m3Sub3(vA[k], m3MulSV3(c->invMassA, P));
            wA[k + 1] = m3Sub3(wA[k], m3MulMV3(c->invIA, m3Cross3(cp->rA, P)));
        }
        if (c->dynB)
        {
            vB[k + 1] = m3Add3(vB[k], m3MulSV3(c->invMassB, P));
            wB[k + 1] = m3Add3(wB[k], m3MulMV3(c->invIB, m3Cross3(cp->rB, P)));
        }
    }
    m3real passNormalBar = 0.0f;
    m3real twistLimitBar = 0.0f;
    if (!useBias)
    {
        // Twist and friction, forward values.
        m3Vec3 vAt = vA[M3_MANIFOLD_MAX_POINTS], wAt = wA[M3_MANIFOLD_MAX_POINTS];
        m3Vec3 vBt = vB[M3_MANIFOLD_MAX_POINTS], wBt = wB[M3_MANIFOLD_MAX_POINTS];
        m3Vec3 wRel = m3Sub3(wBt, wAt);
        m3real ts = m3Dot3(c->normal, wRel);
        m3real maxTwist = c->friction * twistLimit;
        m3real delta0 = -c->twistMass * ts;
        m3real oldT = rec->twist;
        m3real newT = oldT + delta0;
        int tclamp = 0;
        if (newT < -maxTwist)
        {
            newT = -maxTwist;
            tclamp = -1;
        }
        else if (newT > maxTwist)
        {
            newT = maxTwist;
            tclamp = 1;
        }
        m3real tdelta = newT - oldT;
        m3Vec3 tw = m3MulSV3(tdelta, c->normal);
        m3Vec3 wAf = c->dynA ? m3Sub3(wAt, m3MulMV3(c->invIA, tw)) : wAt;
        m3Vec3 wBf = c->dynB ? m3Add3(wBt, m3MulMV3(c->invIB, tw)) : wBt;
        m3Vec3 vrel = m3Sub3(m3Add3(vBt, m3Cross3(wBf, c->originB)), m3Add3(vAt, m3Cross3(wAf, c->originA)));
        m3real vt1 = m3Dot3(vrel, c->t1) - c->tangentVelocity1;
        m3real vt2 = m3Dot3(vrel, c->t2) - c->tangentVelocity2;
        m3real g1 = rec->f1 - (c->frictionK11 * vt1 + c->frictionK12 * vt2);
        m3real g2 = rec->f2 - (c->frictionK12 * vt1 + c->frictionK22 * vt2);
        m3real maxF = c->friction * passNormal;
        m3real mag2 = g1 * g1 + g2 * g2;
        int fclamp = mag2 > maxF * maxF;
        m3real mag = 0.0f, scale = 1.0f;
        m3real f1 = g1, f2 = g2;
        if (fclamp)
        {
            mag = m3Sqrt(mag2);
            scale = mag > 0.0f ? maxF / mag : 0.0f;
            f1 = g1 * scale;
            f2 = g2 * scale;
        }
        m3real d1 = f1 - rec->f1;
        m3real d2 = f2 - rec->f2;
        m3Vec3 I = m3Add3(m3MulSV3(d1, c->t1), m3MulSV3(d2, c->t2));
        // Friction reverse.
        m3Vec3 Ib = m3AdjApply(c, I, c->originA, c->originB, A, B, cb, &cb->originA, &cb->originB);
        m3real d1b = m3Dot3(Ib, c->t1);
        m3real d2b = m3Dot3(Ib, c->t2);
        m3Acc3(&cb->t1, m3MulSV3(d1, Ib));
        m3Acc3(&cb->t2, m3MulSV3(d2, Ib));
        m3real f1b = cb->f1 + d1b;
        m3real f2b = cb->f2 + d2b;
        m3real F1b = -d1b;
        m3real F2b = -d2b;
        m3real g1b = f1b, g2b = f2b, maxFb = 0.0f;
        if (fclamp)
        {
            g1b = f1b * scale;
            g2b = f2b * scale;
            if (mag > 0.0f)
            {
                m3real sb = f1b * g1 + f2b * g2;
                maxFb = sb / mag;
                m3real magb = -sb * maxF / (mag * mag);
                m3real mag2b = magb / (2.0f * mag);
                g1b += 2.0f * g1 * mag2b;
                g2b += 2.0f * g2 * mag2b;
            }
        }
        cb->friction += maxFb * passNormal;
        passNormalBar += maxFb * c->friction;
        F1b += g1b;
        F2b += g2b;
        cb->k11 -= g1b * vt1;
        cb->k12 -= g1b * vt2 + g2b * vt1;
        cb->k22 -= g2b * vt2;
        m3real vt1b = -(g1b * c->frictionK11 + g2b * c->frictionK12);
        m3real vt2b = -(g1b * c->frictionK12 + g2b * c->frictionK22);
        m3Vec3 vrelb = m3Add3(m3MulSV3(vt1b, c->t1), m3MulSV3(vt2b, c->t2));
        m3Acc3(&cb->t1, m3MulSV3(vt1b, vrel));
        m3Acc3(&cb->t2, m3MulSV3(vt2b, vrel));
        m3Acc3(&B->v, vrelb);
        m3Acc3(&A->v, m3Neg3(vrelb));
        m3Acc3(&B->w, m3Cross3(c->originB, vrelb));
        m3Acc3(&cb->originB, m3Cross3(vrelb, wBf));
        m3Acc3(&A->w, m3Neg3(m3Cross3(c->originA, vrelb)));
        m3Acc3(&cb->originA, m3Neg3(m3Cross3(vrelb, wAf)));
        cb->f1 = F1b;
        cb->f2 = F2b;
        // Twist reverse.
        m3Vec3 twb = m3Z3();
        if (c->dynA)
        {
            twb = m3Sub3(twb, m3MulMTV3(c->invIA, A->w));
            m3AccOuter(&cb->invIA, m3Neg3(A->w), tw);
        }
        if (c->dynB)
        {
            twb = m3Add3(twb, m3MulMTV3(c->invIB, B->w));
            m3AccOuter(&cb->invIB, B->w, tw);
        }
        m3real tdb = m3Dot3(twb, c->normal);
        m3Acc3(&cb->normal, m3MulSV3(tdelta, twb));
        m3real newTb = cb->twist + tdb;
        m3real oldTb = -tdb;
        m3real maxTb = 0.0f;
        m3real delta0b = 0.0f;
        if (tclamp < 0)
        {
            maxTb = -newTb;
        }
        else if (tclamp > 0)
        {
            maxTb = newTb;
        }
        else
        {
            oldTb += newTb;
            delta0b = newTb;
        }
        cb->twistMass -= delta0b * ts;
        m3real tsb = -delta0b * c->twistMass;
        m3Acc3(&cb->normal, m3MulSV3(tsb, wRel));
        m3Vec3 wRelb = m3MulSV3(tsb, c->normal);
        m3Acc3(&B->w, wRelb);
        m3Acc3(&A->w, m3Neg3(wRelb));
        cb->friction += maxTb * twistLimit;
        twistLimitBar += maxTb * c->friction;
        cb->twist = oldTb;
    }
    // Normal points in reverse.
#pragma unroll
    for (int32_t k = M3_MANIFOLD_MAX_POINTS - 1; k >= 0; --k)
    {
        if (k >= np)
        {
            continue;
        }
        const m3CPoint* cp = &c->points[k];
        m3real d = deltaK[k];
        m3Vec3 P = m3MulSV3(d, c->normal);
        m3Vec3 Pb = m3AdjApply(c, P, cp->rA, cp->rB, A, B, cb, &cb->rA[k], &cb->rB[k]);
        m3real db = m3Dot3(Pb, c->normal);
        m3Acc3(&cb->normal, m3MulSV3(d, Pb));
        m3real nlb = cb->lambda[k] + passNormalBar + cp->leverArm * twistLimitBar + db;
        cb->leverArm[k] += twistLimitBar * newK[k];
        m3real lamb = -db;
        m3real impb = 0.0f;
        if (activeK[k])
        {
            lamb += nlb;
            impb = nlb;
        }
        m3real vnb = -cp->normalMass * msK[k] * impb;
        m3real biasb = vnb;
        cb->normalMass[k] -= msK[k] * (vnK[k] + biasK[k]) * impb;
        lamb -= isK[k] * impb;
        // vn = ((vB + wB x rB) - (vA + wA x rA)) . n, at the point's input state
        m3Vec3 n = c->normal;
        m3Acc3(&B->v, m3MulSV3(vnb, n));
        m3Acc3(&A->v, m3MulSV3(-vnb, n));
        m3Acc3(&B->w, m3MulSV3(vnb, m3Cross3(cp->rB, n)));
        m3Acc3(&A->w, m3MulSV3(-vnb, m3Cross3(cp->rA, n)));
        m3Acc3(&cb->rB[k], m3MulSV3(vnb, m3Cross3(n, wB[k])));
        m3Acc3(&cb->rA[k], m3MulSV3(-vnb, m3Cross3(n, wA[k])));
        m3Vec3 rel = m3Sub3(m3Add3(vB[k], m3Cross3(wB[k], cp->rB)), m3Add3(vA[k], m3Cross3(wA[k], cp->rA)));
        m3Acc3(&cb->normal, m3MulSV3(vnb, rel));
        m3real sb = 0.0f;
        if (branchK[k] == 0)
        {
            sb = invH * biasb;
        }
        else if (branchK[k] == 1)
        {
            sb = c->softness.biasRate * biasb;
        }
        cb->baseSeparation[k] += sb;
        m3Acc3(&cb->normal, m3MulSV3(sb, dsK[k]));
        m3Vec3 dsb = m3MulSV3(sb, n);
        m3Acc3(&B->dp, dsb);
        m3Acc3(&A->dp, m3Neg3(dsb));
        m3Vec3 rBb = m3AdjRotate(dqB, cp->rB, dsb, &B->dq);
        m3Vec3 rAb = m3AdjRotate(dqA, cp->rA, m3Neg3(dsb), &A->dq);
        m3Acc3(&cb->rB[k], rBb);
        m3Acc3(&cb->rA[k], rAb);
        cb->lambda[k] = lamb;
    }
}

// Reverse of m3WarmStartOne at the accumulators `rec` (the velocity
// updates are additive, so the velocity bars pass through unchanged).
M3_HD void m3AdjWarmStart(const m3Constraint* c, const m3TapePass* rec, const m3BodyBar* A,
                          const m3BodyBar* B, m3ConBar* cb)
{
    for (int32_t k = 0; k < c->pointCount; ++k)
    {
        const m3CPoint* cp = &c->points[k];
        m3Vec3 P = m3MulSV3(rec->lambda[k], c->normal);
        m3Vec3 Pb = m3AdjApply(c, P, cp->rA, cp->rB, A, B, cb, &cb->rA[k], &cb->rB[k]);
        cb->lambda[k] += m3Dot3(Pb, c->normal);
        m3Acc3(&cb->normal, m3MulSV3(rec->lambda[k], Pb));
    }
    m3Vec3 F = m3Add3(m3MulSV3(rec->f1, c->t1), m3MulSV3(rec->f2, c->t2));
    m3Vec3 Fb = m3AdjApply(c, F, c->originA, c->originB, A, B, cb, &cb->originA, &cb->originB);
    cb->f1 += m3Dot3(Fb, c->t1);
    cb->f2 += m3Dot3(Fb, c->t2);
    m3Acc3(&cb->t1, m3MulSV3(rec->f1, Fb));
    m3Acc3(&cb->t2, m3MulSV3(rec->f2, Fb));
    m3Vec3 tw = m3MulSV3(rec->twist, c->normal);
    m3Vec3 twb = m3Z3();
    if (c->dynA)
    {
        twb = m3Sub3(twb, m3MulMTV3(c->invIA, A->w));
        m3AccOuter(&cb->invIA, m3Neg3(A->w), tw);
    }
    if (c->dynB)
    {
        twb = m3Add3(twb, m3MulMTV3(c->invIB, B->w));
        m3AccOuter(&cb->invIB, B->w, tw);
    }
    cb->twist += m3Dot3(twb, c->normal);
    m3Acc3(&cb->normal, m3MulSV3(rec->twist, twb));
}

// ---- integration ------------------------------------------------------------

// Per-body parameter adjoints.
typedef struct m3ParamBar
{
    m3real mass;     // d/d(mass) at fixed shape: inverse mass and inverse inertia scale as 1/m
    m3real inertia;  // sum of Ibar_ij I_ij of the forward inertia (scales as m)
    m3real invI;     // sum of invIbar_ij invI_ij of the local inverse inertia (scales as 1/m)
    m3real invMass;  // adjoint of the inverse mass
    m3Vec3 force;    // applied force, summed over every use
    m3real friction; // the body's shape friction
} m3ParamBar;

// Reverse of m3GyroscopicOmega.  Returns the bar of the input angular
// velocity; adds q̄ and the inertia-scale bar.
M3_HD m3Vec3 m3AdjGyroscopic(m3Mat3 inertia, m3Quat q, m3Vec3 w, m3real h, m3Vec3 yb, m3Quat* qb,
                             m3real* inertiaScaleBar)
{
    const m3real i00 = inertia.cx.x, i01 = inertia.cy.x, i02 = inertia.cz.x;
    const m3real i11 = inertia.cy.y, i12 = inertia.cz.y, i22 = inertia.cz.z;
    if (i01 == 0.0f && i02 == 0.0f && i12 == 0.0f && i00 == i11 && i11 == i22)
    {
        return yb;
    }
    m3Vec3 om = m3InvRotateVec3(q, w);
    const m3real w1 = om.x, w2 = om.y, w3 = om.z;
    const m3real Iw1 = i00 * w1 + i01 * w2 + i02 * w3;
    const m3real Iw2 = i01 * w1 + i11 * w2 + i12 * w3;
    const m3real Iw3 = i02 * w1 + i12 * w2 + i22 * w3;
    m3Vec3 b = m3V3(h * (w2 * Iw3 - w3 * Iw2), h * (w3 * Iw1 - w1 * Iw3), h * (w1 * Iw2 - w2 * Iw1));
    m3Mat3 J;
    J.cx = m3V3(i00 + h * (w2 * i02 - w3 * i01), i01 + h * (w3 * i00 - w1 * i02 - Iw3),
                i02 + h * (w1 * i01 - w2 * i00 + Iw2));
    J.cy = m3V3(i01 + h * (w2 * i12 - w3 * i11 + Iw3), i11 + h * (w3 * i01 - w1 * i12),
                i12 + h * (w1 * i11 - w2 * i01 - Iw1));
    J.cz = m3V3(i02 + h * (w2 * i22 - w3 * i12 - Iw2), i12 + h * (w3 * i02 - w1 * i22 + Iw1),
                i22 + h * (w1 * i12 - w2 * i02));
    m3Vec3 x = m3Solve3(&J, b);
    m3Vec3 om2 = m3Sub3(om, x);
    // out = R om2
    m3Vec3 om2b = m3AdjRotate(q, om2, yb, qb);
    m3Vec3 omb = om2b;
    m3Vec3 xb = m3Neg3(om2b);
    // x = J^-1 b (zero when J is singular)
    m3Mat3 JT;
    JT.cx = m3V3(J.cx.x, J.cy.x, J.cz.x);
    JT.cy = m3V3(J.cx.y, J.cy.y, J.cz.y);
    JT.cz = m3V3(J.cx.z, J.cy.z, J.cz.z);
    m3Vec3 cxy = m3Cross3(J.cy, J.cz);
    m3real det = m3Dot3(J.cx, cxy);
    m3Vec3 bb = m3Z3();
    m3Mat3 Jb = m3ZeroMat3();
    if (det != 0.0f)
    {
        bb = m3Solve3(&JT, xb);
        m3AccOuter(&Jb, m3Neg3(bb), x);
    }
    // b = h (om x Iw)
    m3real I1b = 0.0f, I2b = 0.0f, I3b = 0.0f;
    m3real w1b = 0.0f, w2b = 0.0f, w3b = 0.0f;
    w2b += h * (bb.x * Iw3 - bb.z * Iw1);
    w3b += h * (-bb.x * Iw2 + bb.y * Iw1);
    w1b += h * (-bb.y * Iw3 + bb.z * Iw2);
    I3b += h * (bb.x * w2 - bb.y * w1);
    I2b += h * (-bb.x * w3 + bb.z * w1);
    I1b += h * (bb.y * w3 - bb.z * w2);
    // J entries
    m3real b00 = 0.0f, b01 = 0.0f, b02 = 0.0f, b11 = 0.0f, b12 = 0.0f, b22 = 0.0f;
    m3real g;
    g = Jb.cx.x; b00 += g; b02 += g * h * w2; b01 -= g * h * w3; w2b += g * h * i02; w3b -= g * h * i01;
    g = Jb.cx.y; b01 += g; b00 += g * h * w3; b02 -= g * h * w1; w3b += g * h * i00; w1b -= g * h * i02; I3b -= g * h;
    g = Jb.cx.z; b02 += g; b01 += g * h * w1; b00 -= g * h * w2; w1b += g * h * i01; w2b -= g * h * i00; I2b += g * h;
    g = Jb.cy.x; b01 += g; b12 += g * h * w2; b11 -= g * h * w3; w2b += g * h * i12; w3b -= g * h * i11; I3b += g * h;
    g = Jb.cy.y; b11 += g; b01 += g * h * w3; b12 -= g * h * w1; w3b += g * h * i01; w1b -= g * h * i12;
    g = Jb.cy.z; b12 += g; b11 += g * h * w1; b01 -= g * h * w2; w1b += g * h * i11; w2b -= g * h * i01; I1b -= g * h;
    g = Jb.cz.x; b02 += g; b22 += g * h * w2; b12 -= g * h * w3; w2b += g * h * i22; w3b -= g * h * i12; I2b -= g * h;
    g = Jb.cz.y; b12 += g; b02 += g * h * w3; b22 -= g * h * w1; w3b += g * h * i02; w1b -= g * h * i22; I1b += g * h;
    g = Jb.cz.z; b22 += g; b12 += g * h * w1; b02 -= g * h * w2; w1b += g * h * i12; w2b -= g * h * i02;
    // Iw = I om
    w1b += I1b * i00 + I2b * i01 + I3b * i02;
    w2b += I1b * i01 + I2b * i11 + I3b * i12;
    w3b += I1b * i02 + I2b * i12 + I3b * i22;
    b00 += I1b * w1;
    b01 += I1b * w2 + I2b * w1;
    b02 += I1b * w3 + I3b * w1;
    b11 += I2b * w2;
    b12 += I2b * w3 + I3b * w2;
    b22 += I3b * w3;
    *inertiaScaleBar += b00 * i00 + b01 * i01 + b02 * i02 + b11 * i11 + b12 * i12 + b22 * i22;
    omb = m3Add3(omb, m3V3(w1b, w2b, w3b));
    // om = R^T w
    return m3AdjInvRotate(q, w, omb, qb);
}

// Reverse of m3IntegrateVelocity for a dynamic body (no torque, speed caps
// not reached: both checked by the caller).  bar->v/w hold the output bars
// on entry and the input bars on return.
M3_HD void m3AdjIntegrateVelocity(const m3View* v, int32_t i, m3Transform xf, m3Vec3 lin, m3Vec3 ang,
                                  m3Vec3 force, m3real h, m3BodyBar* bar, m3ParamBar* pb)
{
    m3real dl = 1.0f / (1.0f + h * v->linDamp[i]);
    m3real da = 1.0f / (1.0f + h * v->angDamp[i]);
    m3Vec3 lin2b = m3MulSV3(dl, bar->v);
    pb->invMass += h * m3Dot3(force, lin2b);
    pb->force = m3Add3(pb->force, m3MulSV3(h * v->invMass[i], lin2b));
    bar->v = lin2b;
    m3Vec3 ang1 = m3MulSV3(da, ang);
    m3Vec3 ang1b = m3AdjGyroscopic(v->inertia[i], xf.q, ang1, h, bar->w, &bar->q, &pb->inertia);
    bar->w = m3MulSV3(da, ang1b);
    (void)lin;
}

// Reverse of m3IntegratePosition (no motion locks).  bar holds the output
// bars (pose, deltas; v/w bars of the unchanged velocities) on entry.
M3_HD void m3AdjIntegratePosition(const m3View* v, int32_t i, m3Transform xf, m3Vec3 lin, m3Vec3 ang,
                                  m3Quat dq, m3real h, m3BodyBar* bar)
{
    m3Vec3 lc = v->lc[i];
    m3Vec3 dw = m3MulSV3(h, ang);
    m3Quat q2 = m3IntegrateRotation(xf.q, dw);
    m3Vec3 cb = bar->p;
    m3Quat q2b = bar->q;
    m3AdjRotate(q2, lc, m3Neg3(bar->p), &q2b);
    m3Quat qb = m3ZQ();
    m3Vec3 dwb = m3AdjIntegrateRotation(xf.q, dw, q2b, &qb);
    m3Quat dqb = m3ZQ();
    dwb = m3Add3(dwb, m3AdjIntegrateRotation(dq, dw, bar->dq, &dqb));
    bar->v = m3Add3(bar->v, m3MulSV3(h, bar->dp));
    bar->w = m3Add3(bar->w, m3MulSV3(h, dwb));
    m3AdjRotate(xf.q, lc, cb, &qb);
    bar->v = m3Add3(bar->v, m3MulSV3(h, cb));
    bar->p = cb;
    bar->q = qb;
    bar->dq = dqb;
}

// ---- prepare (m3PrepareContact) ---------------------------------------------

// Adjoints of a manifold's inputs to prepare.
typedef struct m3ManBar
{
    m3Vec3 normal;
    m3Vec3 anchorA[M3_MANIFOLD_MAX_POINTS], anchorB[M3_MANIFOLD_MAX_POINTS];
    m3real separation[M3_MANIFOLD_MAX_POINTS];
    m3real normalImpulse[M3_MANIFOLD_MAX_POINTS];
    m3Vec3 frictionImpulse;
    m3real twistImpulse;
} m3ManBar;

// World inverse inertia R I^-1 R^T: adds q̄ and the local-inverse scale bar.
M3_HD void m3AdjWorldInvInertia(m3Quat q, m3Mat3 il, m3Mat3 Wb, m3Quat* qb, m3real* invIScaleBar)
{
    m3Vec3 e[3] = {m3V3(1.0f, 0.0f, 0.0f), m3V3(0.0f, 1.0f, 0.0f), m3V3(0.0f, 0.0f, 1.0f)};
    m3Vec3 cols[3] = {Wb.cx, Wb.cy, Wb.cz};
    m3Mat3 ilb = m3ZeroMat3();
    for (int j = 0; j < 3; ++j)
    {
        m3Vec3 u = m3InvRotateVec3(q, e[j]);
        m3Vec3 z = m3MulMV3(il, u);
        m3Vec3 zb = m3AdjRotate(q, z, cols[j], qb);
        m3AccOuter(&ilb, zb, u);
        m3Vec3 ub = m3MulMTV3(il, zb);
        m3AdjInvRotate(q, e[j], ub, qb);
    }
    *invIScaleBar += m3FrobDot(ilb, il);
}

M3_HD m3real m3EffMassK(const m3Constraint* c, m3Vec3 rA, m3Vec3 rB, m3Vec3 dir)
{
    m3Vec3 arm1 = m3Cross3(rA, dir);
    m3Vec3 arm2 = m3Cross3(rB, dir);
    return c->invMassA + c->invMassB + m3Dot3(arm1, m3MulMV3(c->invIA, arm1)) +
           m3Dot3(arm2, m3MulMV3(c->invIB, arm2));
}

// k = mA + mB + (rA x d).IA(rA x d) + (rB x d).IB(rB x d): given k̄ adds the
// bars of rA, rB, d and the mass terms.
M3_HD void m3AdjEffK(const m3Constraint* c, m3Vec3 rA, m3Vec3 rB, m3Vec3 dir, m3real kb, m3Vec3* rAb,
                     m3Vec3* rBb, m3Vec3* dirb, m3ConBar* cb)
{
    cb->invMassA += kb;
    cb->invMassB += kb;
    m3Vec3 arm1 = m3Cross3(rA, dir);
    m3Vec3 arm2 = m3Cross3(rB, dir);
    m3Vec3 arm1b = m3MulSV3(kb, m3Add3(m3MulMV3(c->invIA, arm1), m3MulMTV3(c->invIA, arm1)));
    m3Vec3 arm2b = m3MulSV3(kb, m3Add3(m3MulMV3(c->invIB, arm2), m3MulMTV3(c->invIB, arm2)));
    m3AccOuter(&cb->invIA, m3MulSV3(kb, arm1), arm1);
    m3AccOuter(&cb->invIB, m3MulSV3(kb, arm2), arm2);
    m3Acc3(rAb, m3Cross3(dir, arm1b));
    m3Acc3(dirb, m3Cross3(arm1b, rA));
    m3Acc3(rBb, m3Cross3(dir, arm2b));
    m3Acc3(dirb, m3Cross3(arm2b, rB));
}

// Reverse of m3PrepareContact for constraint c of manifold m, given the bars
// of its fields (cb; running accumulator bars are those at the first
// substep's start).  Adds the manifold bars, the bodies' rotation bars
// (through the world inverse inertia) and the parameter bars.
M3_HD void m3AdjPrepare(const m3View* v, const m3Manifold* m, const m3Constraint* c, m3ConBar* cb,
                        m3ManBar* mb, m3Quat* qbA, m3Quat* qbB, m3ParamBar* pA, m3ParamBar* pB,
                        int32_t shapeA, int32_t shapeB)
{
    const int32_t np = c->pointCount;
    m3Vec3 n = c->normal;
    m3Vec3 nb = cb->normal;
    // twist mass
    {
        m3Vec3 iSumN = m3Add3(m3MulMV3(c->invIA, n), m3MulMV3(c->invIB, n));
        m3real twistK = m3Dot3(n, iSumN);
        if (twistK > 0.0f)
        {
            m3real kb = -cb->twistMass / (twistK * twistK);
            m3Acc3(&nb, m3MulSV3(kb, iSumN));
            m3Vec3 iSb = m3MulSV3(kb, n);
            m3Acc3(&nb, m3Add3(m3MulMTV3(c->invIA, iSb), m3MulMTV3(c->invIB, iSb)));
            m3AccOuter(&cb->invIA, iSb, n);
            m3AccOuter(&cb->invIB, iSb, n);
        }
    }
    mb->twistImpulse += cb->twist;
    // friction impulses: F1 = F . t1, F2 = F . t2
    m3Acc3(&mb->frictionImpulse, m3Add3(m3MulSV3(cb->f1, c->t1), m3MulSV3(cb->f2, c->t2)));
    m3Vec3 t1b = m3Add3(cb->t1, m3MulSV3(cb->f1, m->frictionImpulse));
    m3Vec3 t2b = m3Add3(cb->t2, m3MulSV3(cb->f2, m->frictionImpulse));
    // friction K
    m3Vec3 oAb = cb->originA;
    m3Vec3 oBb = cb->originB;
    {
        m3Vec3 rtA1 = m3Cross3(c->originA, c->t1);
        m3Vec3 rtA2 = m3Cross3(c->originA, c->t2);
        m3Vec3 rtB1 = m3Cross3(c->originB, c->t1);
        m3Vec3 rtB2 = m3Cross3(c->originB, c->t2);
        m3real k11 = c->invMassA + c->invMassB + m3Dot3(rtA1, m3MulMV3(c->invIA, rtA1)) +
                     m3Dot3(rtB1, m3MulMV3(c->invIB, rtB1));
        m3real k22 = c->invMassA + c->invMassB + m3Dot3(rtA2, m3MulMV3(c->invIA, rtA2)) +
                     m3Dot3(rtB2, m3MulMV3(c->invIB, rtB2));
        m3real k12 = m3Dot3(rtA1, m3MulMV3(c->invIA, rtA2)) + m3Dot3(rtB1, m3MulMV3(c->invIB, rtB2));
        m3real det = k11 * k22 - k12 * k12;
        if (det > 0.0f)
        {
            m3real invDet = 1.0f / det;
            // K11 = k22 invDet, K22 = k11 invDet, K12 = -k12 invDet
            m3real k22b = cb->k11 * invDet;
            m3real k11b = cb->k22 * invDet;
            m3real k12b = -cb->k12 * invDet;
            m3real invDetb = cb->k11 * k22 + cb->k22 * k11 - cb->k12 * k12;
            m3real detb = -invDetb * invDet * invDet;
            k11b += detb * k22;
            k22b += detb * k11;
            k12b -= detb * 2.0f * k12;
            // k11 = mA + mB + rtA1.IA rtA1 + rtB1.IB rtB1, etc.
            cb->invMassA += k11b + k22b;
            cb->invMassB += k11b + k22b;
            m3Vec3 rtA1b = m3Add3(m3MulSV3(k11b, m3Add3(m3MulMV3(c->invIA, rtA1), m3MulMTV3(c->invIA, rtA1))),
                                  m3MulSV3(k12b, m3MulMV3(c->invIA, rtA2)));
            m3Vec3 rtA2b = m3Add3(m3MulSV3(k22b, m3Add3(m3MulMV3(c->invIA, rtA2), m3MulMTV3(c->invIA, rtA2))),
                                  m3MulSV3(k12b, m3MulMTV3(c->invIA, rtA1)));
            m3Vec3 rtB1b = m3Add3(m3MulSV3(k11b, m3Add3(m3MulMV3(c->invIB, rtB1), m3MulMTV3(c->invIB, rtB1))),
                                  m3MulSV3(k12b, m3MulMV3(c->invIB, rtB2)));
            m3Vec3 rtB2b = m3Add3(m3MulSV3(k22b, m3Add3(m3MulMV3(c->invIB, rtB2), m3MulMTV3(c->invIB, rtB2))),
                                  m3MulSV3(k12b, m3MulMTV3(c->invIB, rtB1)));
            m3AccOuter(&cb->invIA, m3MulSV3(k11b, rtA1), rtA1);
            m3AccOuter(&cb->invIA, m3MulSV3(k22b, rtA2), rtA2);
            m3AccOuter(&cb->invIA, m3MulSV3(k12b, rtA1), rtA2);
            m3AccOuter(&cb->invIB, m3MulSV3(k11b, rtB1), rtB1);
            m3AccOuter(&cb->invIB, m3MulSV3(k22b, rtB2), rtB2);
            m3AccOuter(&cb->invIB, m3MulSV3(k12b, rtB1), rtB2);
            // rt = o x t
            m3Acc3(&oAb, m3Add3(m3Cross3(c->t1, rtA1b), m3Cross3(c->t2, rtA2b)));
            m3Acc3(&t1b, m3Cross3(rtA1b, c->originA));
            m3Acc3(&t2b, m3Cross3(rtA2b, c->originA));
            m3Acc3(&oBb, m3Add3(m3Cross3(c->t1, rtB1b), m3Cross3(c->t2, rtB2b)));
            m3Acc3(&t1b, m3Cross3(rtB1b, c->originB));
            m3Acc3(&t2b, m3Cross3(rtB2b, c->originB));
        }
    }
    // lever arms: |rA_k - oA|
    m3Vec3 rAb[M3_MANIFOLD_MAX_POINTS];
    m3Vec3 rBb[M3_MANIFOLD_MAX_POINTS];
    for (int32_t k = 0; k < np; ++k)
    {
        rAb[k] = cb->rA[k];
        rBb[k] = cb->rB[k];
        m3Vec3 d = m3Sub3(c->points[k].rA, c->originA);
        m3real len = m3Length3(d);
        if (len > 0.0f)
        {
            m3Vec3 db = m3MulSV3(cb->leverArm[k] / len, d);
            m3Acc3(&rAb[k], db);
            oAb = m3Sub3(oAb, db);
        }
    }
    // origins: mean of the anchors
    m3real invCount = 1.0f / (m3real)np;
    for (int32_t k = 0; k < np; ++k)
    {
        m3Acc3(&rAb[k], m3MulSV3(invCount, oAb));
        m3Acc3(&rBb[k], m3MulSV3(invCount, oBb));
    }
    // points
    for (int32_t k = 0; k < np; ++k)
    {
        const m3CPoint* cp = &c->points[k];
        mb->normalImpulse[k] += cb->lambda[k];
        // normalMass = 1 / k
        m3real K = m3EffMassK(c, cp->rA, cp->rB, n);
        if (K > 0.0f)
        {
            m3real kb = -cb->normalMass[k] / (K * K);
            m3AdjEffK(c, cp->rA, cp->rB, n, kb, &rAb[k], &rBb[k], &nb, cb);
        }
        // base = sep - (rB.n - rA.n)
        m3real bb = cb->baseSeparation[k];
        mb->separation[k] += bb;
        m3Acc3(&rBb[k], m3MulSV3(-bb, n));
        m3Acc3(&rAb[k], m3MulSV3(bb, n));
        m3Acc3(&nb, m3MulSV3(-bb, m3Sub3(cp->rB, cp->rA)));
        m3Acc3(&mb->anchorA[k], rAb[k]);
        m3Acc3(&mb->anchorB[k], rBb[k]);
    }
    // tangent basis: t1 = normalize(n x axis), t2 = n x t1
    {
        m3real ax = m3AbsF(n.x), ay = m3AbsF(n.y), az = m3AbsF(n.z);
        m3Vec3 axis;
        if (ax <= ay && ax <= az)
        {
            axis = m3V3(1.0f, 0.0f, 0.0f);
        }
        else if (ay <= az)
        {
            axis = m3V3(0.0f, 1.0f, 0.0f);
        }
        else
        {
            axis = m3V3(0.0f, 0.0f, 1.0f);
        }
        m3Acc3(&nb, m3Cross3(c->t1, t2b));
        m3Vec3 t1tot = m3Add3(t1b, m3Cross3(t2b, n));
        m3Vec3 crossv = m3Cross3(n, axis);
        m3Vec3 cvb = m3AdjNormalize3(crossv, t1tot);
        m3Acc3(&nb, m3Cross3(axis, cvb));
    }
    m3Acc3(&mb->normal, nb);
    // friction = sqrt(fa fb)
    {
        m3real fa = v->friction[shapeA];
        m3real fb = v->friction[shapeB];
        if (c->friction > 0.0f)
        {
            m3real g = cb->friction / (2.0f * c->friction);
            pA->friction += g * fb;
            pB->friction += g * fa;
        }
    }
    // masses
    pA->invMass += cb->invMassA;
    pB->invMass += cb->invMassB;
    if (c->dynA)
    {
        m3AdjWorldInvInertia(v->xf[c->bodyA].q, v->invI[c->bodyA], cb->invIA, qbA, &pA->invI);
    }
    if (c->dynB)
    {
        m3AdjWorldInvInertia(v->xf[c->bodyB].q, v->invI[c->bodyB], cb->invIB, qbB, &pB->invI);
    }
}

// ---- narrowphase geometry ---------------------------------------------------

// Plane against hull (m3CollidePlaneHull): points are hull vertices.  Adds
// the hull body's position and rotation bars.
M3_HD void m3AdjPlaneHull(const m3View* v, int32_t shapeA, int32_t planeShape, int32_t hullShape,
                          const m3Manifold* m, const m3ManBar* mb, m3Vec3* pb, m3Quat* qb)
{
    const m3HullData* hull = &v->hulls[v->hullIndex[hullShape]];
    int32_t hullBody = v->shapeBody[hullShape];
    m3Transform xf = v->xf[hullBody];
    m3Vec3 n = v->geom[planeShape].v;
    m3Vec3 lc = v->lc[hullBody];
    int swapped = planeShape != shapeA;
    for (int32_t k = 0; k < m->pointCount; ++k)
    {
        m3Vec3 vert = hull->vertices[m->points[k].id];
        m3Vec3 aPlane = swapped ? mb->anchorB[k] : mb->anchorA[k];
        m3Vec3 aHull = swapped ? mb->anchorA[k] : mb->anchorB[k];
        m3real sb = mb->separation[k];
        // anchor on the hull = R v - R lc
        m3AdjRotate(xf.q, vert, aHull, qb);
        m3AdjRotate(xf.q, lc, m3Neg3(aHull), qb);
        // separation = n . (p + R v) - offset;  plane anchor = p + R v - n sep - c_plane
        m3Vec3 wb = m3Add3(m3MulSV3(sb, n), m3Sub3(aPlane, m3MulSV3(m3Dot3(n, aPlane), n)));
        m3Acc3(pb, wb);
        m3AdjRotate(xf.q, vert, wb, qb);
    }
}

// Closest points of two segments (m3SegmentClosest) in reverse.  Adds the
// bars of p2 and d2 (segment 1 is constant in the caller's frame).
M3_HD void m3AdjSegmentClosest(m3Vec3 p1, m3Vec3 d1, m3Vec3 p2, m3Vec3 d2, m3Vec3 c1b, m3Vec3 c2b,
                               m3Vec3* p2b, m3Vec3* d2b)
{
    m3Vec3 r = m3Sub3(p1, p2);
    m3real a = m3Dot3(d1, d1);
    m3real e = m3Dot3(d2, d2);
    m3real f = m3Dot3(d2, r);
    m3real c = m3Dot3(d1, r);
    m3real b = m3Dot3(d1, d2);
    m3real denom = a * e - b * b;
    m3real s0raw = denom > 1.0e-9f ? (b * f - c * e) / denom : 0.0f;
    m3real s0 = denom > 1.0e-9f ? m3ClampF(s0raw, 0.0f, 1.0f) : 0.0f;
    m3real traw = e > 1.0e-9f ? (b * s0 + f) / e : 0.0f;
    m3real t = e > 1.0e-9f ? m3ClampF(traw, 0.0f, 1.0f) : 0.0f;
    m3real sraw = a > 1.0e-9f ? (b * t - c) / a : 0.0f;
    m3real s = a > 1.0e-9f ? m3ClampF(sraw, 0.0f, 1.0f) : 0.0f;
    // c1 = p1 + s d1; c2 = p2 + t d2
    m3real sb = m3Dot3(c1b, d1);
    m3Acc3(p2b, c2b);
    m3real tb = m3Dot3(c2b, d2);
    m3Acc3(d2b, m3MulSV3(t, c2b));
    m3real bb = 0.0f, cb = 0.0f, eb = 0.0f, fb = 0.0f;
    // s = clamp((b t - c) / a)
    if (a > 1.0e-9f && sraw > 0.0f && sraw < 1.0f)
    {
        m3real q = sb / a;
        bb += q * t;
        tb += q * b;
        cb -= q;
    }
    m3real s0b = 0.0f;
    if (e > 1.0e-9f && traw > 0.0f && traw < 1.0f)
    {
        m3real q = tb / e;
        bb += q * s0;
        s0b += q * b;
        fb += q;
        eb -= q * traw;
    }
    if (denom > 1.0e-9f && s0raw > 0.0f && s0raw < 1.0f)
    {
        m3real q = s0b / denom;
        bb += q * f;
        fb += q * b;
        cb -= q * e;
        eb -= q * c;
        m3real denomb = -q * s0raw;
        eb += denomb * a;
        bb -= denomb * 2.0f * b;
    }
    // b = d1.d2; f = d2.r; e = d2.d2; c = d1.r; r = p1 - p2
    m3Acc3(d2b, m3Add3(m3MulSV3(bb, d1), m3Add3(m3MulSV3(fb, r), m3MulSV3(2.0f * eb, d2))));
    m3Vec3 rb = m3Add3(m3MulSV3(fb, d2), m3MulSV3(cb, d1));
    m3Acc3(p2b, m3Neg3(rb));
}

#define M3_ADJ_CLIP_PASSES 8
#define M3_ADJ_CLIP_POINTS 16

// How a hull pair's local manifold was lifted to world (m3CollidePairHulls):
// normal = R(qA) n, anchorA = R(qA) aA - R(qA) lcA,
// anchorB = pA + R(qA) aB - pB - R(qB) lcB.  `mb` holds the world bars; the
// lift's own bars add into the body bars.
typedef struct m3HullLift
{
    m3Quat qA, qB;
    m3Vec3 lcA, lcB;
    const m3ManBar* mb;
    m3Quat* qbA;
    m3Quat* qbB;
    m3Vec3* pbA;
    m3Vec3* pbB;
} m3HullLift;

// Bars of one local point from the world bars through the lift.
M3_HD void m3AdjLiftPoint(const m3HullLift* L, int32_t j, m3Vec3 aA, m3Vec3 aB, m3Vec3* aAb, m3Vec3* aBb)
{
    const m3ManBar* mb = L->mb;
    *aAb = m3AdjRotate(L->qA, aA, mb->anchorA[j], L->qbA);
    m3AdjRotate(L->qA, L->lcA, m3Neg3(mb->anchorA[j]), L->qbA);
    *aBb = m3AdjRotate(L->qA, aB, mb->anchorB[j], L->qbA);
    m3Acc3(L->pbA, mb->anchorB[j]);
    m3Acc3(L->pbB, m3Neg3(mb->anchorB[j]));
    m3AdjRotate(L->qB, L->lcB, m3Neg3(mb->anchorB[j]), L->qbB);
}

// Hull against hull in A's frame (m3CollideHullsQ), evaluated once with its
// clip operations kept, then in reverse: the world bars come back through
// the lift to local bars, and from those to the bars of (q, p), B's pose in
// A's frame.  Returns 0 when the clip is larger than this reverse keeps or
// the point count disagrees (the caller counts the pair as unsupported).
M3_HD int m3AdjCollideHulls(const m3HullData* hullA, const m3HullData* hullB, m3Quat q, m3Vec3 p,
                            int32_t pointCount, const m3HullLift* lift, m3Quat* qb, m3Vec3* pb)
{
    m3FaceQuery faceA = m3QueryFaces(hullA, hullB, q, p, 1);
    m3FaceQuery faceB = m3QueryFaces(hullB, hullA, q, p, 0);
    m3EdgeQuery edge = m3QueryEdges(hullA, hullB, q, p);
    const m3real linearSlop = M3_LINEAR_SLOP;
    m3real maxFace = m3MaxF(faceA.separation, faceB.separation);
    if (edge.indexA >= 0 && edge.separation > maxFace + 0.1f * linearSlop)
    {
        // Edge contact: one point from the closest points of the two edges;
        // normal = the edge axis, separation = axis . (pB - pA).
        m3HullHalfEdge eA = hullA->edges[edge.indexA];
        m3HullHalfEdge tA = hullA->edges[edge.indexA + 1];
        m3HullHalfEdge eB = hullB->edges[edge.indexB];
        m3HullHalfEdge tB = hullB->edges[edge.indexB + 1];
        m3Vec3 pA = hullA->vertices[eA.origin];
        m3Vec3 dA = m3Sub3(hullA->vertices[tA.origin], pA);
        m3Vec3 vB0 = hullB->vertices[eB.origin];
        m3Vec3 vB1 = hullB->vertices[tB.origin];
        m3Vec3 pBv = m3Add3(m3RotateVec3(q, vB0), p);
        m3Vec3 qBv = m3Add3(m3RotateVec3(q, vB1), p);
        m3Vec3 dB = m3Sub3(qBv, pBv);
        if (pointCount != 1)
        {
            return 0;
        }
        m3Vec3 cA;
        m3Vec3 cB;
        m3SegmentClosest(pA, dA, pBv, dB, &cA, &cB);
        m3ManBar lbv;
        memset(&lbv, 0, sizeof(lbv));
        m3ManBar* lb = &lbv;
        lb->normal = m3AdjRotate(lift->qA, edge.axis, lift->mb->normal, lift->qbA);
        lb->separation[0] = lift->mb->separation[0];
        m3AdjLiftPoint(lift, 0, cA, cB, &lb->anchorA[0], &lb->anchorB[0]);
        m3Vec3 pBb = m3Z3();
        m3Vec3 dBb = m3Z3();
        m3AdjSegmentClosest(pA, dA, pBv, dB, lb->anchorA[0], lb->anchorB[0], &pBb, &dBb);
        // axis (query form): eB = R vB1 + p - (R vB0 + p) with the edge's own
        // vertex order in the query (twin origin minus origin)
        m3Vec3 qBq = m3Add3(m3RotateVec3(q, hullB->vertices[tB.origin]), p);
        m3Vec3 pBq = m3Add3(m3RotateVec3(q, hullB->vertices[eB.origin]), p);
        m3Vec3 eBq = m3Sub3(qBq, pBq);
        m3Vec3 eAq = dA;
        m3Vec3 raw = m3Cross3(eAq, eBq);
        m3real len = m3Length3(raw);
        m3Vec3 axis = m3MulSV3(1.0f / len, raw);
        int flipped = m3Dot3(axis, m3Sub3(pA, hullA->center)) < 0.0f;
        if (flipped)
        {
            axis = m3Neg3(axis);
        }
        // separation = axis . (pBq - pA); normal = axis
        m3real sepb = lb->separation[0];
        m3Vec3 axisb = m3Add3(lb->normal, m3MulSV3(sepb, m3Sub3(pBq, pA)));
        m3Vec3 pBqb = m3MulSV3(sepb, axis);
        if (flipped)
        {
            axisb = m3Neg3(axisb);
        }
        // axis_unflipped = raw / len
        m3Vec3 rawb = m3AdjNormalize3(raw, axisb);
        m3Vec3 eBqb = m3Cross3(rawb, eAq);
        // eBq = qBq - pBq
        m3Vec3 qBqb = eBqb;
        pBqb = m3Sub3(pBqb, eBqb);
        // pBq, qBq, pBv, qBv are rotated vertices plus p
        m3Acc3(pb, m3Add3(pBqb, qBqb));
        m3AdjRotate(q, hullB->vertices[eB.origin], pBqb, qb);
        m3AdjRotate(q, hullB->vertices[tB.origin], qBqb, qb);
        // dB = qBv - pBv
        m3Vec3 qBvb = dBb;
        m3Vec3 pBvb = m3Sub3(pBb, dBb);
        m3Acc3(pb, m3Add3(pBvb, qBvb));
        m3AdjRotate(q, vB0, pBvb, qb);
        m3AdjRotate(q, vB1, qBvb, qb);
        return 1;
    }
    int refIsA = faceB.separation <= faceA.separation + 0.1f * linearSlop ? 1 : 0;
    const m3HullData* ref = refIsA ? hullA : hullB;
    const m3HullData* inc = refIsA ? hullB : hullA;
    int32_t refFace = refIsA ? faceA.faceIndex : faceB.faceIndex;
    m3Vec3 refN = ref->faceNormals[refFace];
    m3real refOff = ref->faceOffsets[refFace];
    m3Vec3 refNInc = refIsA ? m3InvRotateVec3(q, refN) : m3RotateVec3(q, refN);
    int32_t incFace = 0;
    m3real minDot = 3.4e38f;
    for (int32_t f = 0; f < inc->faceCount; ++f)
    {
        m3real d = m3Dot3(refNInc, inc->faceNormals[f]);
        if (d < minDot)
        {
            minDot = d;
            incFace = f;
        }
    }
    // Forward clip with its operations kept.
    m3Vec3 poly[M3_ADJ_CLIP_PASSES + 1][M3_ADJ_CLIP_POINTS];
    uint16_t polyId[M3_ADJ_CLIP_POINTS];
    uint16_t outId[M3_ADJ_CLIP_POINTS];
    int16_t srcA[M3_ADJ_CLIP_PASSES][M3_ADJ_CLIP_POINTS];
    int16_t srcB[M3_ADJ_CLIP_PASSES][M3_ADJ_CLIP_POINTS]; // -1 for a copy
    m3real tK[M3_ADJ_CLIP_PASSES][M3_ADJ_CLIP_POINTS];
    m3real dcK[M3_ADJ_CLIP_PASSES][M3_ADJ_CLIP_POINTS];
    m3real dnK[M3_ADJ_CLIP_PASSES][M3_ADJ_CLIP_POINTS];
    int32_t countK[M3_ADJ_CLIP_PASSES + 1];
    m3Vec3 sideNK[M3_ADJ_CLIP_PASSES];
    int32_t polyCount = inc->faceVertCounts[incFace];
    if (polyCount > M3_ADJ_CLIP_POINTS)
    {
        return 0;
    }
    m3Vec3 incV[M3_ADJ_CLIP_POINTS];
    for (int32_t k = 0; k < polyCount; ++k)
    {
        uint8_t vi = inc->faceIndices[inc->faceVertStart[incFace] + k];
        m3Vec3 vv = inc->vertices[vi];
        incV[k] = vv;
        poly[0][k] = refIsA ? m3Add3(m3RotateVec3(q, vv), p) : m3InvRotateVec3(q, m3Sub3(vv, p));
        polyId[k] = vi;
    }
    countK[0] = polyCount;
    int32_t n = ref->faceVertCounts[refFace];
    int32_t start = ref->faceVertStart[refFace];
    int32_t passes = 0;
    for (int32_t e = 0; e < n && polyCount > 0; ++e)
    {
        if (passes >= M3_ADJ_CLIP_PASSES)
        {
            return 0;
        }
        m3Vec3 v1 = ref->vertices[ref->faceIndices[start + e]];
        m3Vec3 v2 = ref->vertices[ref->faceIndices[start + (e + 1) % n]];
        m3Vec3 tangent = m3Normalize3(m3Sub3(v2, v1));
        m3Vec3 sideN = m3Cross3(tangent, refN);
        m3real sideOff = m3Dot3(sideN, v1);
        sideNK[passes] = sideN;
        int32_t outCount = 0;
        for (int32_t k = 0; k < polyCount; ++k)
        {
            m3Vec3 cur = poly[passes][k];
            int32_t k1 = (k + 1) % polyCount;
            m3Vec3 nxt = poly[passes][k1];
            m3real dc = m3Dot3(sideN, cur) - sideOff;
            m3real dn = m3Dot3(sideN, nxt) - sideOff;
            if (dc <= 0.0f)
            {
                if (outCount >= M3_ADJ_CLIP_POINTS)
                {
                    return 0;
                }
                poly[passes + 1][outCount] = cur;
                srcA[passes][outCount] = (int16_t)k;
                srcB[passes][outCount] = -1;
                outId[outCount] = polyId[k];
                outCount += 1;
            }
            if (dc * dn < 0.0f)
            {
                if (outCount >= M3_ADJ_CLIP_POINTS)
                {
                    return 0;
                }
                m3real t = dc / (dc - dn);
                poly[passes + 1][outCount] = m3Add3(m3MulSV3(1.0f - t, cur), m3MulSV3(t, nxt));
                srcA[passes][outCount] = (int16_t)k;
                srcB[passes][outCount] = (int16_t)k1;
                tK[passes][outCount] = t;
                dcK[passes][outCount] = dc;
                dnK[passes][outCount] = dn;
                outId[outCount] = (uint16_t)(0x4000u | ((uint32_t)e << 8) | polyId[k]);
                outCount += 1;
            }
        }
        for (int32_t k = 0; k < outCount; ++k)
        {
            polyId[k] = outId[k];
        }
        polyCount = outCount;
        passes += 1;
        countK[passes] = outCount;
    }
    // Match the manifold's points to the final polygon by id.
    m3Vec3 polyb[M3_ADJ_CLIP_POINTS];
    for (int32_t k = 0; k < polyCount; ++k)
    {
        polyb[k] = m3Z3();
    }
    m3Vec3 outN = refIsA ? refN : m3Neg3(m3RotateVec3(q, refN));
    m3ManBar lbv;
    memset(&lbv, 0, sizeof(lbv));
    m3ManBar* lb = &lbv;
    lb->normal = m3AdjRotate(lift->qA, outN, lift->mb->normal, lift->qbA);
    if (!refIsA)
    {
        // normal = -R(q) refN
        m3AdjRotate(q, refN, m3Neg3(lb->normal), qb);
    }
    // The forward's selection: points below the margin, the deepest four
    // (ties to the lower polygon id), emitted in ascending id order.
    int32_t candIdx[M3_ADJ_CLIP_POINTS];
    m3real candSep[M3_ADJ_CLIP_POINTS];
    int32_t candCount = 0;
    for (int32_t k = 0; k < polyCount; ++k)
    {
        m3real sep = m3Dot3(refN, poly[passes][k]) - refOff;
        if (sep < M3_SPECULATIVE_DISTANCE)
        {
            candIdx[candCount] = k;
            candSep[candCount] = sep;
            candCount += 1;
        }
    }
    int32_t want = candCount < M3_MANIFOLD_MAX_POINTS ? candCount : M3_MANIFOLD_MAX_POINTS;
    int32_t kept[M3_MANIFOLD_MAX_POINTS];
    uint32_t used = 0u;
    for (int32_t k = 0; k < want; ++k)
    {
        int32_t best = -1;
        for (int32_t c = 0; c < candCount; ++c)
        {
            if ((used >> c) & 1u)
            {
                continue;
            }
            if (best < 0 || candSep[c] < candSep[best] ||
                (candSep[c] == candSep[best] && polyId[candIdx[c]] < polyId[candIdx[best]]))
            {
                best = c;
            }
        }
        used |= 1u << best;
        kept[k] = best;
    }
    for (int32_t a = 0; a < want; ++a)
    {
        for (int32_t b = a + 1; b < want; ++b)
        {
            if (polyId[candIdx[kept[b]]] < polyId[candIdx[kept[a]]])
            {
                int32_t tmp = kept[a];
                kept[a] = kept[b];
                kept[b] = tmp;
            }
        }
    }
    if (want != pointCount)
    {
        return 0;
    }
    for (int32_t j = 0; j < pointCount; ++j)
    {
        int32_t kk = candIdx[kept[j]];
        m3Vec3 onInc = poly[passes][kk];
        m3real sep = m3Dot3(refN, onInc) - refOff;
        m3Vec3 onRef = m3Sub3(onInc, m3MulSV3(sep, refN));
        m3Vec3 wInc = refIsA ? onInc : m3Add3(m3RotateVec3(q, onInc), p);
        m3Vec3 wRef = refIsA ? onRef : m3Add3(m3RotateVec3(q, onRef), p);
        lb->separation[j] = lift->mb->separation[j];
        m3AdjLiftPoint(lift, j, refIsA ? wRef : wInc, refIsA ? wInc : wRef, &lb->anchorA[j], &lb->anchorB[j]);
        m3Vec3 wIncb = refIsA ? lb->anchorB[j] : lb->anchorA[j];
        m3Vec3 wRefb = refIsA ? lb->anchorA[j] : lb->anchorB[j];
        m3Vec3 onIncb, onRefb;
        if (refIsA)
        {
            onIncb = wIncb;
            onRefb = wRefb;
        }
        else
        {
            m3Acc3(pb, m3Add3(wIncb, wRefb));
            onIncb = m3AdjRotate(q, onInc, wIncb, qb);
            onRefb = m3AdjRotate(q, onRef, wRefb, qb);
        }
        m3real sepb = lb->separation[j] - m3Dot3(onRefb, refN);
        onIncb = m3Add3(onIncb, m3Add3(onRefb, m3MulSV3(sepb, refN)));
        m3Acc3(&polyb[kk], onIncb);
    }
    // Clip passes in reverse.
    for (int32_t ps = passes - 1; ps >= 0; --ps)
    {
        m3Vec3 inb[M3_ADJ_CLIP_POINTS];
        int32_t inCount = countK[ps];
        for (int32_t k = 0; k < inCount; ++k)
        {
            inb[k] = m3Z3();
        }
        for (int32_t o = 0; o < countK[ps + 1]; ++o)
        {
            int32_t a = srcA[ps][o];
            int32_t b = srcB[ps][o];
            if (b < 0)
            {
                m3Acc3(&inb[a], polyb[o]);
                continue;
            }
            m3real t = tK[ps][o];
            m3Vec3 cur = poly[ps][a];
            m3Vec3 nxt = poly[ps][b];
            m3Vec3 yb = polyb[o];
            m3Acc3(&inb[a], m3MulSV3(1.0f - t, yb));
            m3Acc3(&inb[b], m3MulSV3(t, yb));
            m3real tb = m3Dot3(yb, m3Sub3(nxt, cur));
            m3real dc = dcK[ps][o];
            m3real dn = dnK[ps][o];
            m3real den = (dc - dn) * (dc - dn);
            m3real dcb = tb * (-dn) / den;
            m3real dnb = tb * dc / den;
            m3Acc3(&inb[a], m3MulSV3(dcb, sideNK[ps]));
            m3Acc3(&inb[b], m3MulSV3(dnb, sideNK[ps]));
        }
        for (int32_t k = 0; k < inCount; ++k)
        {
            polyb[k] = inb[k];
        }
    }
    // Incident vertices into the reference frame.
    for (int32_t k = 0; k < countK[0]; ++k)
    {
        if (refIsA)
        {
            m3Acc3(pb, polyb[k]);
            m3AdjRotate(q, incV[k], polyb[k], qb);
        }
        else
        {
            m3Vec3 xb = m3AdjInvRotate(q, m3Sub3(incV[k], p), polyb[k], qb);
            m3Acc3(pb, m3Neg3(xb));
        }
    }
    return 1;
}

// Hull pair in world (m3CollidePairHulls) in reverse: adds both bodies'
// position and rotation bars.
M3_HD int m3AdjPairHulls(const m3View* v, int32_t shapeA, int32_t shapeB, const m3Manifold* m,
                         const m3ManBar* mb, m3Vec3* pbA, m3Quat* qbA, m3Vec3* pbB, m3Quat* qbB)
{
    int32_t bodyA = v->shapeBody[shapeA];
    int32_t bodyB = v->shapeBody[shapeB];
    m3Transform xfA;
    m3Quat qRel;
    m3Vec3 pRel;
    m3HullPairFrame(v, shapeA, shapeB, &xfA, &qRel, &pRel);
    m3Transform xfB = m3ShapeXf(v, shapeB);
    const m3HullData* hA = &v->hulls[v->hullIndex[shapeA]];
    const m3HullData* hB = &v->hulls[v->hullIndex[shapeB]];
    m3HullLift lift;
    lift.qA = xfA.q;
    lift.qB = xfB.q;
    lift.lcA = v->lc[bodyA];
    lift.lcB = v->lc[bodyB];
    lift.mb = mb;
    lift.qbA = qbA;
    lift.qbB = qbB;
    lift.pbA = pbA;
    lift.pbB = pbB;
    m3Quat qRelb = m3ZQ();
    m3Vec3 pRelb = m3Z3();
    if (!m3AdjCollideHulls(hA, hB, qRel, pRel, m->pointCount, &lift, &qRelb, &pRelb))
    {
        return 0;
    }
    // qRel = conj(qA) qB
    m3Quat cAb = m3ZQ();
    m3AdjMulQuat(m3Conj(xfA.q), xfB.q, qRelb, &cAb, qbB);
    qbA->x -= cAb.x;
    qbA->y -= cAb.y;
    qbA->z -= cAb.z;
    qbA->w += cAb.w;
    // pRel = R(qA)^T (pB - pA)
    m3Vec3 dp = m3V3((m3real)(xfB.p.x - xfA.p.x), (m3real)(xfB.p.y - xfA.p.y), (m3real)(xfB.p.z - xfA.p.z));
    m3Vec3 dpb = m3AdjInvRotate(xfA.q, dp, pRelb, qbA);
    m3Acc3(pbB, dpb);
    m3Acc3(pbA, m3Neg3(dpb));
    return 1;
}

// ---- continuous pull-back (m3ContinuousBody, plane targets) ------------------

#define M3_ADJ_TOI_ITERS 25

// Reverse of a pull-back of body i (one shape, planes only): the pose
// bars in `bar` (after the pull-back) become bars of the pose before it
// (xfEnd); bars of the step-start pose (the sweep's start) add into p0b/q0b.
// The time of impact is differentiated through the conservative-advance
// iterations that produced it.  Returns 0 when the pull-back is outside
// this reverse's scope.
M3_HD int m3AdjContinuous(const m3View* v, int32_t i, m3Transform xf0, m3Transform xfEnd, m3BodyBar* bar,
                          m3Vec3* p0b, m3Quat* q0b)
{
    int32_t shape = v->shapeHead[i];
    if (shape < 0 || v->shapeNext[shape] != -1 || v->shapeType[shape] != M3_HULL)
    {
        return 0;
    }
    const m3HullData* hull = &v->hulls[v->hullIndex[shape]];
    m3Vec3 lc = v->lc[i];
    // Sweep (m3RelativeSweep with base = com0): c1 = 0.
    m3Vec3 rlc0 = m3RotateVec3(xf0.q, lc);
    m3Pos3 base;
    base.x = xf0.p.x + (double)rlc0.x;
    base.y = xf0.p.y + (double)rlc0.y;
    base.z = xf0.p.z + (double)rlc0.z;
    m3Sweep sw;
    sw.localCenter = lc;
    sw.c1 = m3V3(0.0f, 0.0f, 0.0f);
    m3Vec3 rlcE = m3RotateVec3(xfEnd.q, lc);
    sw.c2 = m3V3((m3real)(xfEnd.p.x + (double)rlcE.x - base.x), (m3real)(xfEnd.p.y + (double)rlcE.y - base.y),
                 (m3real)(xfEnd.p.z + (double)rlcE.z - base.z));
    sw.q1 = xf0.q;
    sw.q2 = xfEnd.q;
    // Replay the plane passes; keep the iterations of the one that set the fraction.
    m3real fraction = 1.0f;
    int32_t bestPlane = -1;
    m3real tIt[M3_ADJ_TOI_ITERS + 1];
    int32_t kIt[M3_ADJ_TOI_ITERS + 1];
    m3real sepIt[M3_ADJ_TOI_ITERS + 1];
    int32_t nIt = 0;
    m3real target = 0.0f, rate = 0.0f;
    for (int32_t pi = 0; pi < v->planeCount; ++pi)
    {
        int32_t ps = v->planes[pi];
        m3Vec3 n = v->geom[ps].v;
        m3real offset = v->geom[ps].s - (m3real)((double)n.x * base.x + (double)n.y * base.y + (double)n.z * base.z);
        const m3real linearSlop = M3_LINEAR_SLOP;
        m3real tgt = m3MaxF(linearSlop, 0.0f - linearSlop);
        m3real tolerance = 0.25f * linearSlop;
        m3Vec3 travel = m3Sub3(sw.c2, sw.c1);
        m3Quat dq = m3MulQuat(sw.q2, m3Q(-sw.q1.x, -sw.q1.y, -sw.q1.z, sw.q1.w));
        m3real arc = 2.0f * m3Sqrt(dq.x * dq.x + dq.y * dq.y + dq.z * dq.z) * v->maxExt[i];
        m3real rt = m3Sqrt(m3Dot3(travel, travel)) + arc;
        if (!(rt > 0.0f))
        {
            continue;
        }
        m3real t = 0.0f;
        m3real tl[M3_ADJ_TOI_ITERS + 1];
        int32_t kl[M3_ADJ_TOI_ITERS + 1];
        m3real sl[M3_ADJ_TOI_ITERS + 1];
        for (int32_t iter = 0; iter < M3_ADJ_TOI_ITERS; ++iter)
        {
            m3Transform xf = m3SweepTransform(&sw, t);
            m3real minD = 3.4e38f;
            int32_t arg = 0;
            for (int32_t k = 0; k < hull->vertexCount; ++k)
            {
                m3Vec3 r = m3RotateVec3(xf.q, hull->vertices[k]);
                m3real d = n.x * ((m3real)xf.p.x + r.x) + n.y * ((m3real)xf.p.y + r.y) + n.z * ((m3real)xf.p.z + r.z);
                if (d < minD)
                {
                    minD = d;
                    arg = k;
                }
            }
            m3real sep = minD - offset;
            tl[iter] = t;
            kl[iter] = arg;
            sl[iter] = sep;
            if (sep <= 0.0f)
            {
                break;
            }
            if (sep <= tgt + tolerance)
            {
                if (t > 0.0f && t < fraction)
                {
                    fraction = t;
                    bestPlane = pi;
                    nIt = iter;
                    for (int32_t q = 0; q <= iter; ++q)
                    {
                        tIt[q] = tl[q];
                        kIt[q] = kl[q];
                        sepIt[q] = sl[q];
                    }
                    target = tgt;
                    rate = rt;
                }
                break;
            }
            t += (sep - tgt) / rt;
            if (t >= fraction)
            {
                break;
            }
        }
    }
    if (bestPlane < 0)
    {
        return 0;
    }
    m3Vec3 n = v->geom[v->planes[bestPlane]].v;
    m3real ts = fraction;
    // Pulled pose: q = normalize(q1 + t (q2s - q1)), p = base + t c2 - R(q) lc.
    m3Quat q2s = sw.q2;
    m3real sign = 1.0f;
    if (sw.q1.x * q2s.x + sw.q1.y * q2s.y + sw.q1.z * q2s.z + sw.q1.w * q2s.w < 0.0f)
    {
        q2s = m3Q(-q2s.x, -q2s.y, -q2s.z, -q2s.w);
        sign = -1.0f;
    }
    m3Vec3 c2b = m3Z3();
    m3Quat q1b = m3ZQ();
    m3Quat q2sb = m3ZQ();
    m3Vec3 baseb = bar->p;
    m3real tb = 0.0f;
    {
        m3Quat qr = m3Q(sw.q1.x + ts * (q2s.x - sw.q1.x), sw.q1.y + ts * (q2s.y - sw.q1.y),
                        sw.q1.z + ts * (q2s.z - sw.q1.z), sw.q1.w + ts * (q2s.w - sw.q1.w));
        m3Quat qo = m3NormalizeQuat(qr);
        m3Quat qob = bar->q;
        m3AdjRotate(qo, lc, m3Neg3(bar->p), &qob);
        m3Vec3 cb = bar->p;
        tb += m3Dot3(cb, m3Sub3(sw.c2, sw.c1));
        c2b = m3Add3(c2b, m3MulSV3(ts, cb));
        m3Quat qrb = m3AdjNormalizeQuat(qr, qob);
        q1b = m3AddQ(q1b, m3Q((1.0f - ts) * qrb.x, (1.0f - ts) * qrb.y, (1.0f - ts) * qrb.z, (1.0f - ts) * qrb.w));
        q2sb = m3AddQ(q2sb, m3Q(ts * qrb.x, ts * qrb.y, ts * qrb.z, ts * qrb.w));
        tb += qrb.x * (q2s.x - sw.q1.x) + qrb.y * (q2s.y - sw.q1.y) + qrb.z * (q2s.z - sw.q1.z) +
              qrb.w * (q2s.w - sw.q1.w);
    }
    // t_n = sum of (sep_i - target) / rate over i < n
    m3real rateb = 0.0f;
    for (int32_t it = nIt - 1; it >= 0; --it)
    {
        m3real sepb = tb / rate;
        rateb -= tb * (sepIt[it] - target) / (rate * rate);
        // sep_i = n . (P(t_i) + R(Q(t_i)) v_k) - offset, offset = s - n . base
        m3real ti = tIt[it];
        m3Quat qr = m3Q(sw.q1.x + ti * (q2s.x - sw.q1.x), sw.q1.y + ti * (q2s.y - sw.q1.y),
                        sw.q1.z + ti * (q2s.z - sw.q1.z), sw.q1.w + ti * (q2s.w - sw.q1.w));
        m3Quat qi = m3NormalizeQuat(qr);
        m3Vec3 Pb = m3MulSV3(sepb, n);
        m3Quat qib = m3ZQ();
        m3AdjRotate(qi, hull->vertices[kIt[it]], Pb, &qib);
        // P = c(t) - R(Q) lc, c(t) = t c2
        m3AdjRotate(qi, lc, m3Neg3(Pb), &qib);
        m3real tib = m3Dot3(Pb, m3Sub3(sw.c2, sw.c1));
        c2b = m3Add3(c2b, m3MulSV3(ti, Pb));
        m3Quat qrb = m3AdjNormalizeQuat(qr, qib);
        q1b = m3AddQ(q1b, m3Q((1.0f - ti) * qrb.x, (1.0f - ti) * qrb.y, (1.0f - ti) * qrb.z, (1.0f - ti) * qrb.w));
        q2sb = m3AddQ(q2sb, m3Q(ti * qrb.x, ti * qrb.y, ti * qrb.z, ti * qrb.w));
        tib += qrb.x * (q2s.x - sw.q1.x) + qrb.y * (q2s.y - sw.q1.y) + qrb.z * (q2s.z - sw.q1.z) +
               qrb.w * (q2s.w - sw.q1.w);
        baseb = m3Add3(baseb, m3MulSV3(sepb, n));
        tb += tib;
    }
    // rate = |travel| + 2 |dq.xyz| maxExt, travel = c2, dq = q2 conj(q1)
    {
        m3Vec3 travel = m3Sub3(sw.c2, sw.c1);
        m3real tl = m3Length3(travel);
        if (tl > 0.0f)
        {
            c2b = m3Add3(c2b, m3MulSV3(rateb / tl, travel));
        }
        m3Quat dq = m3MulQuat(sw.q2, m3Q(-sw.q1.x, -sw.q1.y, -sw.q1.z, sw.q1.w));
        m3real dl = m3Sqrt(dq.x * dq.x + dq.y * dq.y + dq.z * dq.z);
        if (dl > 0.0f)
        {
            m3real g = rateb * 2.0f * v->maxExt[i] / dl;
            m3Quat dqb = m3Q(g * dq.x, g * dq.y, g * dq.z, 0.0f);
            m3Quat q2b = m3ZQ();
            m3Quat cq1b = m3ZQ();
            m3AdjMulQuat(sw.q2, m3Q(-sw.q1.x, -sw.q1.y, -sw.q1.z, sw.q1.w), dqb, &q2b, &cq1b);
            q1b = m3AddQ(q1b, m3Q(-cq1b.x, -cq1b.y, -cq1b.z, cq1b.w));
            q2sb = m3AddQ(q2sb, m3Q(sign * q2b.x, sign * q2b.y, sign * q2b.z, sign * q2b.w));
        }
    }
    // q2s = sign q2; c2 = pEnd + R(qEnd) lc - base; q1 = q0; base = p0 + R(q0) lc
    m3Quat qEb = m3Q(sign * q2sb.x, sign * q2sb.y, sign * q2sb.z, sign * q2sb.w);
    m3AdjRotate(xfEnd.q, lc, c2b, &qEb);
    bar->p = c2b;
    bar->q = qEb;
    baseb = m3Sub3(baseb, c2b);
    m3Acc3(p0b, baseb);
    m3AdjRotate(xf0.q, lc, baseb, q0b);
    m3AccQ(q0b, q1b);
    return 1;
}

// A constraint's contributions to its bodies' bars from prepare and the
// narrowphase, gathered per body afterwards (no two lanes add into one body).
typedef struct m3ConContrib
{
    m3Vec3 pA, pB;
    m3Quat qA, qB;
    m3ParamBar parA, parB;
} m3ConContrib;

// Reverse-pass context of one world.
typedef struct m3AdjWorld
{
    m3BodyBar* body;       // [bodyCap]
    m3ParamBar* param;     // [bodyCap]
    m3ConBar* cons;        // [pairCap] scratch
    m3ManBar* store;       // [pairCap] bars of this step's stored impulses (by pair)
    m3ManBar* storePrev;   // [pairCap] bars of the previous step's stored impulses
    m3Constraint* prep;    // [pairCap] scratch: constraints re-prepared
    m3ConContrib* contrib; // [pairCap] scratch
    m3Vec3* ccdP;          // [bodyCap] scratch: step-start pose bars of pull-backs
    m3Quat* ccdQ;
    int32_t unsupported; // pairs or bodies outside the reverse's scope
    int32_t ccdPulled;   // pulled-back bodies passed through
} m3AdjWorld;

M3_HD void m3AccParam(m3ParamBar* a, const m3ParamBar* b)
{
    a->mass += b->mass;
    a->inertia += b->inertia;
    a->invI += b->invI;
    a->invMass += b->invMass;
    a->force = m3Add3(a->force, b->force);
    a->friction += b->friction;
}

// Reverse of one constraint's solve in pass `mode` of substep `sub`, with
// lane-private bars standing in for static sides.
M3_HD void m3AdjSolveSlot(const m3View* v, const m3Tape* tape, m3AdjWorld* aw, int32_t sub, int32_t mode,
                          int32_t k, m3real invH, m3real pushMax)
{
    const m3Constraint* c = &aw->prep[k];
    const m3TapePass* rec = &tape->pass[((int64_t)sub * 2 + (mode - 1)) * tape->pairCap + k];
    const m3Vec3* dpArr = mode == 2 ? tape->rdp : tape->sdp;
    const m3Quat* dqArr = mode == 2 ? tape->rdq : tape->sdq;
    const int32_t base = sub * tape->bodyCap;
    m3BodyBar dummyA, dummyB;
    memset(&dummyA, 0, sizeof(dummyA));
    memset(&dummyB, 0, sizeof(dummyB));
    m3BodyBar* A = c->dynA ? &aw->body[c->bodyA] : &dummyA;
    m3BodyBar* B = c->dynB ? &aw->body[c->bodyB] : &dummyB;
    m3AdjSolveContact(c, rec, dpArr[base + c->bodyA], dqArr[base + c->bodyA], dpArr[base + c->bodyB],
                      dqArr[base + c->bodyB], invH, pushMax, mode == 1, A, B, &aw->cons[k]);
    (void)v;
}

// Reverse of the step recorded in `tape`, phase by phase.  Lane `lane` of
// `lanes` takes every lanes-th item of each phase and `sync` separates the
// phases (a workgroup barrier on the device; nothing for one lane).  `v`
// is the world's view with its pose, velocities and pairs pointed at the
// tape's (m3AdjTapeView).  On entry aw->body holds the bars of the step's
// final state and aw->store those of the impulses it stored; on return
// aw->body holds the bars of its initial state and aw->storePrev those of
// the previous step's stored impulses.
// The world's view with the step's recorded pose, velocities and pairs.
M3_HD void m3AdjTapeView(m3View* v, const m3Tape* tape)
{
    v->xf = tape->xf0;
    v->lv = tape->lv0;
    v->av = tape->av0;
    v->keys = tape->keys;
    v->manifolds = tape->man;
}

template <typename Sync>
M3_HD void m3AdjStepPar(const m3View* v, const m3SerialParams* p, const m3Tape* tape, m3AdjWorld* aw,
                        int32_t lane, int32_t lanes, Sync sync)
{
    const int32_t nb = v->maxBody;
    const int32_t pairCount = tape->counts[0];
    const int32_t nc = tape->counts[1];
    m3real h = p->dt / (m3real)p->substeps;
    m3real invH = h > 0.0f ? 1.0f / h : 0.0f;
    m3Softness soft = m3MakeSoft(p->contactHertz, p->contactDampingRatio, h);
    m3Softness staticSoft = m3MakeSoft(2.0f * p->contactHertz, p->contactDampingRatio, h);
    const int32_t* cs = tape->colorStarts;
    // storePrev is indexed by the previous step's pairs: clear it whole.
    for (int32_t i = lane; i < tape->pairCap; i += lanes)
    {
        memset(&aw->storePrev[i], 0, sizeof(m3ManBar));
    }
    // Sleep froze velocities to zero; the continuous pass pulled poses back
    // along the step's sweep.
    for (int32_t b = lane; b < nb; b += lanes)
    {
        if (tape->slept[b])
        {
            aw->body[b].v = m3Z3();
            aw->body[b].w = m3Z3();
        }
        aw->ccdP[b] = m3Z3();
        aw->ccdQ[b] = m3ZQ();
        if (tape->ccd[b])
        {
            if (!m3AdjContinuous(v, b, tape->xf0[b], tape->xfEnd[b], &aw->body[b], &aw->ccdP[b], &aw->ccdQ[b]))
            {
                aw->ccdPulled += 1;
            }
        }
        aw->body[b].dp = m3Z3();
        aw->body[b].dq = m3ZQ();
    }
    // Re-prepare and seed the constraint bars from the stored impulses.
    for (int32_t k = lane; k < nc; k += lanes)
    {
        int32_t i = tape->consPair[k];
        m3PrepareContact(v, i, &aw->prep[k], soft, staticSoft);
        m3ConBar* cb = &aw->cons[k];
        memset(cb, 0, sizeof(*cb));
        const m3Constraint* c = &aw->prep[k];
        const m3ManBar* sb = &aw->store[i];
        const m3TapePass* fin = &tape->fin[k];
        for (int32_t q = 0; q < c->pointCount; ++q)
        {
            cb->lambda[q] = sb->normalImpulse[q];
        }
        // stored friction = f1 t1 + f2 t2
        cb->f1 = m3Dot3(sb->frictionImpulse, c->t1);
        cb->f2 = m3Dot3(sb->frictionImpulse, c->t2);
        cb->t1 = m3MulSV3(fin->f1, sb->frictionImpulse);
        cb->t2 = m3MulSV3(fin->f2, sb->frictionImpulse);
        cb->twist = sb->twistImpulse;
        if (c->rollingResistance > 0.0f || c->restitution > 0.0f)
        {
            aw->unsupported += 1;
        }
    }
    sync();
    const int32_t B = tape->bodyCap;
    for (int32_t sub = p->substeps - 1; sub >= 0; --sub)
    {
        const int32_t base = sub * B;
        for (int32_t mode = 2; mode >= 1; --mode)
        {
            // Overflow color first (serial in the forward), then 15..0.
            if (lane == 0)
            {
                for (int32_t k = cs[M3_GRAPH_COLORS + 1] - 1; k >= cs[M3_GRAPH_COLORS]; --k)
                {
                    m3AdjSolveSlot(v, tape, aw, sub, mode, k, invH, p->contactPushMaxSpeed);
                }
            }
            sync();
            for (int32_t color = M3_GRAPH_COLORS - 1; color >= 0; --color)
            {
                if (cs[color + 1] == cs[color])
                {
                    continue;
                }
                for (int32_t k = cs[color] + lane; k < cs[color + 1]; k += lanes)
                {
                    m3AdjSolveSlot(v, tape, aw, sub, mode, k, invH, p->contactPushMaxSpeed);
                }
                sync();
            }
            if (mode == 2)
            {
                for (int32_t b = lane; b < nb; b += lanes)
                {
                    if (!tape->mover[b])
                    {
                        continue;
                    }
                    m3AdjIntegratePosition(v, b, tape->sxf[base + b], tape->plv[base + b],
                                           tape->pav[base + b], tape->sdq[base + b], h, &aw->body[b]);
                }
                sync();
            }
        }
        for (int32_t k = lane; k < nc; k += lanes)
        {
            const m3Constraint* c = &aw->prep[k];
            const m3TapePass* rec = &tape->pass[((int64_t)sub * 2) * tape->pairCap + k];
            m3AdjWarmStart(c, rec, &aw->body[c->bodyA], &aw->body[c->bodyB], &aw->cons[k]);
        }
        sync();
        for (int32_t b = lane; b < nb; b += lanes)
        {
            if (!tape->mover[b] || v->type[b] != M3_DYNAMIC)
            {
                continue;
            }
            m3AdjIntegrateVelocity(v, b, tape->sxf[base + b], tape->slv[base + b], tape->sav[base + b],
                                   tape->force0[b], h, &aw->body[b], &aw->param[b]);
        }
        sync();
    }
    // Prepare and narrowphase in reverse, per constraint.
    for (int32_t k = lane; k < nc; k += lanes)
    {
        int32_t i = tape->consPair[k];
        const m3Constraint* c = &aw->prep[k];
        const m3Manifold* m = &tape->man[i];
        uint64_t key = tape->keys[i];
        int32_t sA = (int32_t)(key >> 32);
        int32_t sB = (int32_t)(key & 0xFFFFFFFFu);
        m3ConContrib* o = &aw->contrib[k];
        memset(o, 0, sizeof(*o));
        m3ManBar mb;
        memset(&mb, 0, sizeof(mb));
        m3AdjPrepare(v, m, c, &aw->cons[k], &mb, &o->qA, &o->qB, &o->parA, &o->parB, sA, sB);
        uint8_t tA = v->shapeType[sA];
        uint8_t tB = v->shapeType[sB];
        if ((tA == M3_PLANE) != (tB == M3_PLANE) && (tA == M3_HULL || tB == M3_HULL))
        {
            int32_t plane = tA == M3_PLANE ? sA : sB;
            int32_t hull = tA == M3_PLANE ? sB : sA;
            int hullIsA = hull == sA;
            m3AdjPlaneHull(v, sA, plane, hull, m, &mb, hullIsA ? &o->pA : &o->pB, hullIsA ? &o->qA : &o->qB);
        }
        else if (tA == M3_HULL && tB == M3_HULL)
        {
            if (!m3AdjPairHulls(v, sA, sB, m, &mb, &o->pA, &o->qA, &o->pB, &o->qB))
            {
                aw->unsupported += 1;
            }
        }
        else
        {
            aw->unsupported += 1;
        }
        // Carried impulses flow to the previous step's stored values (pair
        // keys are unique, so no two pairs share a previous pair).
        int32_t prev = tape->carryPair[i];
        if (prev >= 0)
        {
            m3ManBar* pbar = &aw->storePrev[prev];
            m3Acc3(&pbar->frictionImpulse, mb.frictionImpulse);
            pbar->twistImpulse += mb.twistImpulse;
            for (int32_t q = 0; q < m->pointCount; ++q)
            {
                int32_t o2 = tape->carryPoint[4 * i + q];
                if (o2 >= 0)
                {
                    pbar->normalImpulse[o2] += mb.normalImpulse[q];
                }
            }
        }
    }
    // Pairs that touched but were not solved (both sides asleep) keep their
    // stored impulses through the carry unchanged.
    for (int32_t i = lane; i < pairCount; i += lanes)
    {
        int32_t prev = tape->carryPair[i];
        const m3Manifold* m = &tape->man[i];
        if (prev < 0 || m->pointCount == 0)
        {
            continue;
        }
        int solved = 0;
        for (int32_t k = 0; k < nc; ++k)
        {
            if (tape->consPair[k] == i)
            {
                solved = 1;
                break;
            }
        }
        if (solved)
        {
            continue;
        }
        m3ManBar* pbar = &aw->storePrev[prev];
        const m3ManBar* sb = &aw->store[i];
        m3Acc3(&pbar->frictionImpulse, sb->frictionImpulse);
        pbar->twistImpulse += sb->twistImpulse;
        for (int32_t q = 0; q < m->pointCount; ++q)
        {
            int32_t o2 = tape->carryPoint[4 * i + q];
            if (o2 >= 0)
            {
                pbar->normalImpulse[o2] += sb->normalImpulse[q];
            }
        }
    }
    sync();
    // Gather each body's contributions in constraint order.
    for (int32_t b = lane; b < nb; b += lanes)
    {
        m3BodyBar* bar = &aw->body[b];
        for (int32_t k = 0; k < nc; ++k)
        {
            const m3Constraint* c = &aw->prep[k];
            const m3ConContrib* o = &aw->contrib[k];
            if (c->bodyA == b)
            {
                m3Acc3(&bar->p, o->pA);
                m3AccQ(&bar->q, o->qA);
                m3AccParam(&aw->param[b], &o->parA);
            }
            if (c->bodyB == b)
            {
                m3Acc3(&bar->p, o->pB);
                m3AccQ(&bar->q, o->qB);
                m3AccParam(&aw->param[b], &o->parB);
            }
        }
        m3Acc3(&bar->p, aw->ccdP[b]);
        m3AccQ(&bar->q, aw->ccdQ[b]);
    }
    sync();
}

#if defined(__HIPCC__)
#define M3_HD_MEMBER __host__ __device__
#else
#define M3_HD_MEMBER
#endif

typedef struct m3NoSync
{
    M3_HD_MEMBER void operator()() const {}
} m3NoSync;

// The step reverse on one thread.
M3_HD void m3AdjStep(const m3View* vin, const m3SerialParams* p, const m3Tape* tape, m3AdjWorld* aw)
{
    m3View v = *vin;
    m3AdjTapeView(&v, tape);
    m3AdjStepPar(&v, p, tape, aw, 0, 1, m3NoSync());
}

#endif // M3_ADJOINT_H
