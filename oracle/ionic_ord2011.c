/*
 * Oracle: O'Hara-Rudy 2011 human ventricular model, endocardial cell type
 * (celltype = 0), one step, FP64.  TEST INFRASTRUCTURE ONLY.
 *
 * Follows models_library/Ohara_Rudy/Ohara_Rudy_2011.c:
 *   celltype = 0                     :5
 *   initial state                    :25-65 (V0 = -85.f)
 *   step = rDY copied into sv        :92-107
 *   right-hand side                  :110-691
 * The reference updates in place and in statement order: each gate is advanced
 * by Rush-Larsen from V_n and the currents are then built from the updated
 * gates (:208-395); V is advanced by forward Euler (:570); CaMKt, the
 * concentrations and the release fluxes follow (:572-645).  This restatement
 * keeps that order, the operand order of each expression and the float-rounded
 * literals, so that it agrees bit for bit with the reference compiled by gcc.
 * The celltype branches are resolved for celltype = 0.
 */
#include <math.h>
#include <stdint.h>

#define ORD_NEQ 41

void orc_ord2011_init_cell(double *y) {
    static const double ic[ORD_NEQ] = {
        -85.f, 7, 7, 145, 145, 1.0e-4, 1.0e-4, 1.2, 1.2, 0, 1, 1, 1, 1, 1, 0, 1, 1, 0, 1, 1,
        0,     1, 1, 0,   1,   1,   1,      1,      1,   0,   1, 1, 0, 0, 0, 0, 1, 0, 0, 0};
    for(int s = 0; s < ORD_NEQ; s++) y[s] = ic[s];
}

/* One step: y (in/out), stim in uA/uF (subtracted from dV/dt as in :570). */
void orc_ord2011_step(double *y, double stim, double dt) {
    double v = y[0], nai = y[1], nass = y[2], ki = y[3], kss = y[4], cai = y[5], cass = y[6];
    double cansr = y[7], cajsr = y[8], m = y[9], hf = y[10], hs = y[11], j = y[12], hsp = y[13];
    double jp = y[14], mL = y[15], hL = y[16], hLp = y[17], a = y[18], iF = y[19], iS = y[20];
    double ap = y[21], iFp = y[22], iSp = y[23], d = y[24], ff = y[25], fs = y[26], fcaf = y[27];
    double fcas = y[28], jca = y[29], nca = y[30], ffp = y[31], fcafp = y[32], xrf = y[33];
    double xrs = y[34], xs1 = y[35], xs2 = y[36], xk1 = y[37], Jrelnp = y[38], Jrelp = y[39];
    double CaMKt = y[40];

    /* extracellular ions, buffers, CaMK, physical constants, geometry (:154-195) */
    const double nao = 140.0, cao = 1.8, ko = 5.4;
    const double BSRmax = 0.047, KmBSR = 0.00087, BSLmax = 1.124, KmBSL = 0.0087;
    const double cmdnmax = 0.05, kmcmdn = 0.00238, trpnmax = 0.07, kmtrpn = 0.0005;
    const double csqnmax = 10.0f, kmcsqn = 0.8;
    const double aCaMK = 0.05, bCaMK = 0.00068, CaMKo = 0.05, KmCaM = 0.0015, KmCaMK = 0.15;
    const double R = 8314.0, T = 310.0, F = 96485.0;
    const double L = 0.01, rad = 0.0011;
    const double vcell = 1000.0f * 3.14f * rad * rad * L;
    const double Ageo = 2.0f * 3.14f * rad * rad + 2.0f * 3.14f * rad * L;
    const double Acap = 2 * Ageo;
    const double vmyo = 0.68f * vcell;
    const double vnsr = 0.0552f * vcell;
    const double vjsr = 0.0048f * vcell;
    const double vss = 0.02f * vcell;

    /* reversal potentials and CaMK (:198-206) */
    const double ENa = (R * T / F) * log(nao / nai);
    const double EK = (R * T / F) * log(ko / ki);
    const double EKs = (R * T / F) * log((ko + 0.01833f * nao) / (ki + 0.01833f * nai));
    double CaMKb = CaMKo * (1.0f - CaMKt) / (1.0f + KmCaM / cass);
    double CaMKa = CaMKb + CaMKt;
    const double vffrt = v * F * F / (R * T);
    const double vfrt = v * F / (R * T);

    /* INa (:208-230) */
    const double mss = 1.0f / (1.0f + exp((-(v + 39.57f)) / 9.871f));
    const double tm = 1.0f / (6.765f * exp((v + 11.64f) / 34.77f) + 8.552f * exp(-(v + 77.42f) / 5.955f));
    m = mss - (mss - m) * exp(-dt / tm);
    const double hss = 1.0f / (1 + exp((v + 82.90f) / 6.086f));
    const double thf = 1.0f / (1.432e-5f * exp(-(v + 1.196f) / 6.285f) + 6.149f * exp((v + 0.5096f) / 20.27f));
    const double ths = 1.0f / (0.009794f * exp(-(v + 17.95f) / 28.05f) + 0.3343f * exp((v + 5.730f) / 56.66f));
    const double Ahf = 0.99;
    const double Ahs = 1.0f - Ahf;
    hf = hss - (hss - hf) * exp(-dt / thf);
    hs = hss - (hss - hs) * exp(-dt / ths);
    const double h = Ahf * hf + Ahs * hs;
    const double jss = hss;
    const double tj = 2.038f + 1.0f / (0.02136f * exp(-(v + 100.6f) / 8.281f) + 0.3052f * exp((v + 0.9941f) / 38.45f));
    j = jss - (jss - j) * exp(-dt / tj);
    const double hssp = 1.0f / (1.0f + exp((v + 89.1f) / 6.086f));
    const double thsp = 3.0f * ths;
    hsp = hssp - (hssp - hsp) * exp(-dt / thsp);
    const double hp = Ahf * hf + Ahs * hsp;
    const double tjp = 1.46f * tj;
    jp = jss - (jss - jp) * exp(-dt / tjp);
    const double GNa = 75;
    const double fINap = (1.0f / (1.0f + KmCaMK / CaMKa));
    const double INa = GNa * (v - ENa) * m * m * m * ((1.0f - fINap) * h * j + fINap * hp * jp);

    /* INaL (:232-248) */
    const double mLss = 1.0f / (1.0f + exp((-(v + 42.85f)) / 5.264f));
    const double tmL = tm;
    mL = mLss - (mLss - mL) * exp(-dt / tmL);
    const double hLss = 1.0f / (1.0f + exp((v + 87.61f) / 7.488f));
    const double thL = 200.0;
    hL = hLss - (hLss - hL) * exp(-dt / thL);
    const double hLssp = 1.0f / (1.0f + exp((v + 93.81f) / 7.488f));
    const double thLp = 3.0f * thL;
    hLp = hLssp - (hLssp - hLp) * exp(-dt / thLp);
    const double GNaL = 0.0075;
    const double fINaLp = (1.0f / (1.0f + KmCaMK / CaMKa));
    const double INaL = GNaL * (v - ENa) * mL * ((1.0f - fINaLp) * hL + fINaLp * hLp);

    /* Ito (:250-293) */
    const double ass = 1.0f / (1.0f + exp((-(v - 14.34f)) / 14.82f));
    const double ta = 1.0515f / (1.0f / (1.2089f * (1.0f + exp(-(v - 18.4099f) / 29.3814f))) + 3.5f / (1.0f + exp((v + 100.0f) / 29.3814f)));
    a = ass - (ass - a) * exp(-dt / ta);
    const double iss = 1.0f / (1.0f + exp((v + 43.94f) / 5.711f));
    const double delta_epi = 1.0f;
    double tiF = 4.562f + 1.0f / (0.3933f * exp((-(v + 100.0f)) / 100.0f) + 0.08004f * exp((v + 50.0f) / 16.59f));
    double tiS = 23.62f + 1.0f / (0.001416f * exp((-(v + 96.52f)) / 59.05f) + 1.780e-8f * exp((v + 114.1f) / 8.079f));
    tiF *= delta_epi;
    tiS *= delta_epi;
    const double AiF = 1.0f / (1.0f + exp((v - 213.6f) / 151.2f));
    const double AiS = 1.0f - AiF;
    iF = iss - (iss - iF) * exp(-dt / tiF);
    iS = iss - (iss - iS) * exp(-dt / tiS);
    const double ii = AiF * iF + AiS * iS;
    const double assp = 1.0f / (1.0f + exp((-(v - 24.34f)) / 14.82f));
    ap = assp - (assp - ap) * exp(-dt / ta);
    const double dti_develop = 1.354f + 1.0e-4f / (exp((v - 167.4f) / 15.89f) + exp(-(v - 12.23f) / 0.2154f));
    const double dti_recover = 1.0f - 0.5f / (1.0f + exp((v + 70.0f) / 20.0f));
    const double tiFp = dti_develop * dti_recover * tiF;
    const double tiSp = dti_develop * dti_recover * tiS;
    iFp = iss - (iss - iFp) * exp(-dt / tiFp);
    iSp = iss - (iss - iSp) * exp(-dt / tiSp);
    const double ip = AiF * iFp + AiS * iSp;
    const double Gto = 0.02;
    const double fItop = (1.0f / (1.0f + KmCaMK / CaMKa));
    const double Ito = Gto * (v - EK) * ((1.0f - fItop) * a * ii + fItop * ap * ip);

    /* ICaL, ICaNa, ICaK (:295-350) */
    const double dss = 1.0f / (1.0f + exp((-(v + 3.940f)) / 4.230f));
    const double td = 0.6f + 1.0f / (exp(-0.05f * (v + 6.0f)) + exp(0.09f * (v + 14.0f)));
    d = dss - (dss - d) * exp(-dt / td);
    const double fss = 1.0f / (1.0f + exp((v + 19.58f) / 3.696f));
    const double tff = 7.0f + 1.0f / (0.0045f * exp(-(v + 20.0f) / 10.0f) + 0.0045f * exp((v + 20.0f) / 10.0f));
    const double tfs = 1000.0f + 1.0f / (0.000035f * exp(-(v + 5.0f) / 4.0f) + 0.000035f * exp((v + 5.0f) / 6.0f));
    const double Aff = 0.6;
    const double Afs = 1.0f - Aff;
    ff = fss - (fss - ff) * exp(-dt / tff);
    fs = fss - (fss - fs) * exp(-dt / tfs);
    const double f = Aff * ff + Afs * fs;
    const double fcass = fss;
    const double tfcaf = 7.0f + 1.0f / (0.04f * exp(-(v - 4.0f) / 7.0f) + 0.04f * exp((v - 4.0f) / 7.0f));
    const double tfcas = 100.0f + 1.0f / (0.00012f * exp(-v / 3.0f) + 0.00012f * exp(v / 7.0f));
    const double Afcaf = 0.3f + 0.6f / (1.0f + exp((v - 10.0f) / 10.0f));
    const double Afcas = 1.0f - Afcaf;
    fcaf = fcass - (fcass - fcaf) * exp(-dt / tfcaf);
    fcas = fcass - (fcass - fcas) * exp(-dt / tfcas);
    const double fca = Afcaf * fcaf + Afcas * fcas;
    const double tjca = 75.0;
    jca = fcass - (fcass - jca) * exp(-dt / tjca);
    const double tffp = 2.5f * tff;
    ffp = fss - (fss - ffp) * exp(-dt / tffp);
    const double fp = Aff * ffp + Afs * fs;
    const double tfcafp = 2.5f * tfcaf;
    fcafp = fcass - (fcass - fcafp) * exp(-dt / tfcafp);
    const double fcap = Afcaf * fcafp + Afcas * fcas;
    const double Kmn = 0.002;
    const double k2n = 1000.0;
    const double km2n = jca * 1.0f;
    const double anca = 1.0f / (k2n / km2n + pow(1.0f + Kmn / cass, 4.0));
    nca = anca * k2n / km2n - (anca * k2n / km2n - nca) * exp(-km2n * dt);
    const double PhiCaL = 4.0f * vffrt * (cass * exp(2.0f * vfrt) - 0.341f * cao) / (exp(2.0f * vfrt) - 1.0f);
    const double PhiCaNa = 1.0f * vffrt * (0.75f * nass * exp(1.0f * vfrt) - 0.75f * nao) / (exp(1.0f * vfrt) - 1.0f);
    const double PhiCaK = 1.0f * vffrt * (0.75f * kss * exp(1.0f * vfrt) - 0.75f * ko) / (exp(1.0f * vfrt) - 1.0f);
    const double zca = 2.0f;
    const double PCa = 0.0001;
    const double PCap = 1.1f * PCa;
    const double PCaNa = 0.00125f * PCa;
    const double PCaK = 3.574e-4f * PCa;
    const double PCaNap = 0.00125f * PCap;
    const double PCaKp = 3.574e-4f * PCap;
    const double fICaLp = (1.0f / (1.0f + KmCaMK / CaMKa));
    const double ICaL = (1.0f - fICaLp) * PCa * PhiCaL * d * (f * (1.0f - nca) + jca * fca * nca) +
                        fICaLp * PCap * PhiCaL * d * (fp * (1.0f - nca) + jca * fcap * nca);
    const double ICaNa = (1.0f - fICaLp) * PCaNa * PhiCaNa * d * (f * (1.0f - nca) + jca * fca * nca) +
                         fICaLp * PCaNap * PhiCaNa * d * (fp * (1.0f - nca) + jca * fcap * nca);
    const double ICaK = (1.0f - fICaLp) * PCaK * PhiCaK * d * (f * (1.0f - nca) + jca * fca * nca) +
                        fICaLp * PCaKp * PhiCaK * d * (fp * (1.0f - nca) + jca * fcap * nca);

    /* IKr (:352-369) */
    const double xrss = 1.0f / (1.0f + exp((-(v + 8.337f)) / 6.789f));
    const double txrf = 12.98f + 1.0f / (0.3652f * exp((v - 31.66f) / 3.869f) + 4.123e-5f * exp((-(v - 47.78f)) / 20.38f));
    const double txrs = 1.865f + 1.0f / (0.06629f * exp((v - 34.70f) / 7.355f) + 1.128e-5f * exp((-(v - 29.74f)) / 25.94f));
    const double Axrf = 1.0f / (1.0f + exp((v + 54.81f) / 38.21f));
    const double Axrs = 1.0f - Axrf;
    xrf = xrss - (xrss - xrf) * exp(-dt / txrf);
    xrs = xrss - (xrss - xrs) * exp(-dt / txrs);
    const double xr = Axrf * xrf + Axrs * xrs;
    const double rkr = 1.0f / (1.0f + exp((v + 55.0f) / 75.0f)) * 1.0f / (1.0f + exp((v - 10.0f) / 30.0f));
    const double GKr = 0.046;
    const double IKr = GKr * sqrt(ko / 5.4f) * xr * rkr * (v - EK);

    /* IKs (:371-382) */
    const double xs1ss = 1.0f / (1.0f + exp((-(v + 11.60f)) / 8.932f));
    const double txs1 = 817.3f + 1.0f / (2.326e-4f * exp((v + 48.28f) / 17.80f) + 0.001292f * exp((-(v + 210.0f)) / 230.0f));
    xs1 = xs1ss - (xs1ss - xs1) * exp(-dt / txs1);
    const double xs2ss = xs1ss;
    const double txs2 = 1.0f / (0.01f * exp((v - 50.0f) / 20.0f) + 0.0193f * exp((-(v + 66.54f)) / 31.0f));
    xs2 = xs2ss - (xs2ss - xs2) * exp(-dt / txs2);
    const double KsCa = 1.0f + 0.6f / (1.0f + pow(3.8e-5f / cai, 1.4f));
    const double GKs = 0.0034;
    const double IKs = GKs * KsCa * xs1 * xs2 * (v - EKs);

    /* IK1 (:384-397) */
    const double xk1ss = 1.0f / (1.0f + exp(-(v + 2.5538f * ko + 144.59f) / (1.5692f * ko + 3.8115f)));
    const double txk1 = 122.2f / (exp((-(v + 127.2f)) / 20.36f) + exp((v + 236.8f) / 69.33f));
    xk1 = xk1ss - (xk1ss - xk1) * exp(-dt / txk1);
    const double rk1 = 1.0f / (1.0f + exp((v + 105.8f - 2.6f * ko) / 9.493f));
    const double GK1 = 0.1908;
    const double IK1 = GK1 * sqrt(ko) * rk1 * xk1 * (v - EK);

    /* INaCa, cytosolic then subspace part (:399-495) */
    const double kna1 = 15.0, kna2 = 5.0, kna3 = 88.12, kasymm = 12.5, wna = 6.0e4, wca = 6.0e4;
    const double wnaca = 5.0e3, kcaon = 1.5e6, kcaoff = 5.0e3, qna = 0.5224, qca = 0.1670;
    const double hca = exp((qca * v * F) / (R * T));
    const double hna = exp((qna * v * F) / (R * T));
    const double zna = 1.0f;
    const double Gncx = 0.0008;
    double INaCa_part[2];
    for(int sub = 0; sub < 2; sub++) {
        /* sub 0: bulk cytosol (nai, cai); sub 1: junctional subspace (nass, cass) */
        const double na = sub == 0 ? nai : nass;
        const double ca = sub == 0 ? cai : cass;
        const double h1 = 1 + na / kna3 * (1 + hna);
        const double h2 = (na * hna) / (kna3 * h1);
        const double h3 = 1.0f / h1;
        const double h4 = 1.0f + na / kna1 * (1 + na / kna2);
        const double h5 = na * na / (h4 * kna1 * kna2);
        const double h6 = 1.0f / h4;
        const double h7 = 1.0f + nao / kna3 * (1.0f + 1.0f / hna);
        const double h8 = nao / (kna3 * hna * h7);
        const double h9 = 1.0f / h7;
        /* the reference writes (1.0f + nao/kna2) for the cytosol and (1 + nao/kna2)
         * for the subspace; both are the same double */
        const double h10 = kasymm + 1.0f + nao / kna1 * (1.0f + nao / kna2);
        const double h11 = nao * nao / (h10 * kna1 * kna2);
        const double h12 = 1.0f / h10;
        const double k1 = h12 * cao * kcaon;
        const double k2 = kcaoff;
        const double k3p = h9 * wca;
        const double k3pp = h8 * wnaca;
        const double k3 = k3p + k3pp;
        const double k4p = h3 * wca / hca;
        const double k4pp = h2 * wnaca;
        const double k4 = k4p + k4pp;
        const double k5 = kcaoff;
        const double k6 = h6 * ca * kcaon;
        const double k7 = h5 * h2 * wna;
        const double k8 = h8 * h11 * wna;
        const double x1 = k2 * k4 * (k7 + k6) + k5 * k7 * (k2 + k3);
        const double x2 = k1 * k7 * (k4 + k5) + k4 * k6 * (k1 + k8);
        const double x3 = k1 * k3 * (k7 + k6) + k8 * k6 * (k2 + k3);
        const double x4 = k2 * k8 * (k4 + k5) + k3 * k5 * (k1 + k8);
        const double E1 = x1 / (x1 + x2 + x3 + x4);
        const double E2 = x2 / (x1 + x2 + x3 + x4);
        const double E3 = x3 / (x1 + x2 + x3 + x4);
        const double E4 = x4 / (x1 + x2 + x3 + x4);
        const double KmCaAct = 150.0e-6;
        const double allo = 1.0f / (1.0f + pow(KmCaAct / ca, 2.0f));
        const double JncxNa = 3.0f * (E4 * k7 - E1 * k8) + E3 * k4pp - E2 * k3pp;
        const double JncxCa = E2 * k2 - E1 * k1;
        INaCa_part[sub] = (sub == 0 ? 0.8f : 0.2f) * Gncx * allo * (zna * JncxNa + zca * JncxCa);
    }
    const double INaCa_i = INaCa_part[0];
    const double INaCa_ss = INaCa_part[1];
    const double INaCa = INaCa_i + INaCa_ss;

    /* INaK (:497-545) */
    const double k1p = 949.5, k1m = 182.4, k2p = 687.2, k2m = 39.4, k3p = 1899.0, k3m = 79300.0;
    const double k4p = 639.0, k4m = 40.0, Knai0 = 9.073, Knao0 = 27.78, delta = -0.1550f;
    const double Knai = Knai0 * exp((delta * v * F) / (3.0f * R * T));
    const double Knao = Knao0 * exp(((1.0f - delta) * v * F) / (3.0f * R * T));
    const double Kki = 0.5, Kko = 0.3582, MgADP = 0.05, MgATP = 9.8, Kmgatp = 1.698e-7, H = 1.0e-7;
    const double eP = 4.2, Khp = 1.698e-7, Knap = 224.0, Kxkur = 292.0;
    const double P = eP / (1.0f + H / Khp + nai / Knap + ki / Kxkur);
    const double a1 = (k1p * pow(nai / Knai, 3.0)) / (pow(1.0f + nai / Knai, 3.0) + pow(1.0f + ki / Kki, 2.0f) - 1.0f);
    const double b1 = k1m * MgADP;
    const double a2 = k2p;
    const double b2 = (k2m * pow(nao / Knao, 3.0)) / (pow(1.0f + nao / Knao, 3.0) + pow(1.0f + ko / Kko, 2.0f) - 1.0f);
    const double a3 = (k3p * pow(ko / Kko, 2.0f)) / (pow(1.0f + nao / Knao, 3.0) + pow(1.0f + ko / Kko, 2.0f) - 1.0f);
    const double b3 = (k3m * P * H) / (1.0f + MgATP / Kmgatp);
    const double a4 = (k4p * MgATP / Kmgatp) / (1.0f + MgATP / Kmgatp);
    const double b4 = (k4m * pow(ki / Kki, 2.0f)) / (pow(1.0f + nai / Knai, 3.0) + pow(1.0f + ki / Kki, 2.0f) - 1.0f);
    const double x1 = a4 * a1 * a2 + b2 * b4 * b3 + a2 * b4 * b3 + b3 * a1 * a2;
    const double x2 = b2 * b1 * b4 + a1 * a2 * a3 + a3 * b1 * b4 + a2 * a3 * b4;
    const double x3 = a2 * a3 * a4 + b3 * b2 * b1 + b2 * b1 * a4 + a3 * a4 * b1;
    const double x4 = b4 * b3 * b2 + a3 * a4 * a1 + b2 * a4 * a1 + b3 * b2 * a1;
    const double E1 = x1 / (x1 + x2 + x3 + x4);
    const double E2 = x2 / (x1 + x2 + x3 + x4);
    const double E3 = x3 / (x1 + x2 + x3 + x4);
    const double E4 = x4 / (x1 + x2 + x3 + x4);
    const double zk = 1.0f;
    const double JnakNa = 3.0f * (E1 * a3 - E2 * b3);
    const double JnakK = 2.0f * (E4 * b1 - E3 * a1);
    const double Pnak = 30;
    const double INaK = Pnak * (zna * JnakNa + zk * JnakK);

    /* background currents and pump (:547-566) */
    const double xkb = 1.0f / (1.0f + exp(-(v - 14.48f) / 18.34f));
    const double GKb = 0.003;
    const double IKb = GKb * xkb * (v - EK);
    const double PNab = 3.75e-10;
    const double INab = PNab * vffrt * (nai * exp(vfrt) - nao) / (exp(vfrt) - 1.0f);
    const double PCab = 2.5e-8;
    const double ICab = PCab * 4.0f * vffrt * (cai * exp(2.0f * vfrt) - 0.341f * cao) / (exp(2.0f * vfrt) - 1.0f);
    const double GpCa = 0.0005;
    const double IpCa = GpCa * cai / (0.0005f + cai);

    /* membrane potential (:570) */
    v += -dt * (INa + INaL + Ito + ICaL + ICaNa + ICaK + IKr + IKs + IK1 + INaCa + INaK + INab + IKb + IpCa + ICab + stim);

    /* CaMK (:572-574) */
    CaMKb = CaMKo * (1.0f - CaMKt) / (1.0f + KmCaM / cass);
    CaMKa = CaMKb + CaMKt;
    CaMKt += dt * (aCaMK * CaMKb * (CaMKb + CaMKt) - bCaMK * CaMKt);

    /* diffusion and SR release fluxes (:576-608) */
    const double JdiffNa = (nass - nai) / 2.0f;
    const double JdiffK = (kss - ki) / 2.0f;
    const double Jdiff = (cass - cai) / 0.2f;
    const double bt = 4.75;
    const double a_rel = 0.5f * bt;
    const double Jrel_inf = a_rel * (-ICaL) / (1.0f + pow(1.5f / cajsr, 8.0));
    double tau_rel = bt / (1.0f + 0.0123f / cajsr);
    if(tau_rel < 0.005) tau_rel = 0.005;
    Jrelnp = Jrel_inf - (Jrel_inf - Jrelnp) * exp(-dt / tau_rel);
    const double btp = 1.25f * bt;
    const double a_relp = 0.5f * btp;
    const double Jrel_infp = a_relp * (-ICaL) / (1.0f + pow(1.5f / cajsr, 8.0f));
    double tau_relp = btp / (1.0f + 0.0123f / cajsr);
    if(tau_relp < 0.005f) tau_relp = 0.005f;
    Jrelp = Jrel_infp - (Jrel_infp - Jrelp) * exp(-dt / tau_relp);
    const double fJrelp = (1.0f / (1.0f + KmCaMK / CaMKa));
    const double Jrel = (1.0f - fJrelp) * Jrelnp + fJrelp * Jrelp;

    /* SERCA, leak, transfer (:610-623) */
    const double Jupnp = 0.004375f * cai / (cai + 0.00092f);
    const double Jupp = 2.75f * 0.004375f * cai / (cai + 0.00092f - 0.00017f);
    const double fJupp = (1.0f / (1.0f + KmCaMK / CaMKa));
    const double Jleak = 0.0039375f * cansr / 15.0f;
    const double Jup = (1.0f - fJupp) * Jupnp + fJupp * Jupp - Jleak;
    const double Jtr = (cansr - cajsr) / 100.0f;

    /* concentrations (:625-645) */
    nai += dt * (-(INa + INaL + 3.0 * INaCa_i + 3.0 * INaK + INab) * Acap / (F * vmyo) + JdiffNa * vss / vmyo);
    nass += dt * (-(ICaNa + 3.0 * INaCa_ss) * Acap / (F * vss) - JdiffNa);
    ki += dt * (-(Ito + IKr + IKs + IK1 + IKb + stim - 2.0f * INaK) * Acap / (F * vmyo) + JdiffK * vss / vmyo);
    kss += dt * (-(ICaK)*Acap / (F * vss) - JdiffK);
    const double Bcai = 1.0f / (1.0f + cmdnmax * kmcmdn / pow(kmcmdn + cai, 2.0f) + trpnmax * kmtrpn / pow(kmtrpn + cai, 2.0f));
    cai += dt * (Bcai * (-(IpCa + ICab - 2.0f * INaCa_i) * Acap / (2.0f * F * vmyo) - Jup * vnsr / vmyo + Jdiff * vss / vmyo));
    const double Bcass = 1.0f / (1.0f + BSRmax * KmBSR / pow(KmBSR + cass, 2.0f) + BSLmax * KmBSL / pow(KmBSL + cass, 2.0f));
    cass += dt * (Bcass * (-(ICaL - 2.0f * INaCa_ss) * Acap / (2.0f * F * vss) + Jrel * vjsr / vss - Jdiff));
    cansr += dt * (Jup - Jtr * vjsr / vnsr);
    const double Bcajsr = 1.0f / (1.0f + csqnmax * kmcsqn / pow(kmcsqn + cajsr, 2.0f));
    cajsr += dt * (Bcajsr * (Jtr - Jrel));

    y[0] = v;      y[1] = nai;     y[2] = nass;    y[3] = ki;      y[4] = kss;
    y[5] = cai;    y[6] = cass;    y[7] = cansr;   y[8] = cajsr;   y[9] = m;
    y[10] = hf;    y[11] = hs;     y[12] = j;      y[13] = hsp;    y[14] = jp;
    y[15] = mL;    y[16] = hL;     y[17] = hLp;    y[18] = a;      y[19] = iF;
    y[20] = iS;    y[21] = ap;     y[22] = iFp;    y[23] = iSp;    y[24] = d;
    y[25] = ff;    y[26] = fs;     y[27] = fcaf;   y[28] = fcas;   y[29] = jca;
    y[30] = nca;   y[31] = ffp;    y[32] = fcafp;  y[33] = xrf;    y[34] = xrs;
    y[35] = xs1;   y[36] = xs2;    y[37] = xk1;    y[38] = Jrelnp; y[39] = Jrelp;
    y[40] = CaMKt;
}
