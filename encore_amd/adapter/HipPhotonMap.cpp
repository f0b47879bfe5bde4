// HipPhotonMap.cpp -- see HipPhotonMap.h.
#include "HipPhotonMap.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>

HipPhotonMap::HipPhotonMap(size_t maxPhotonCount, int ordinal, rq_photon_layout layout)
    : m_max(maxPhotonCount), m_caustic(false), m_built(false), m_ordinal(ordinal), m_layout(layout), m_dev(0),
      m_map(0), m_stage(0), m_stageBytes(0) {
    m_photons.reserve(maxPhotonCount);
}

HipPhotonMap::~HipPhotonMap() {
    if (m_stage) (void)hipFree(m_stage);
    if (m_map) rq_photons_destroy(m_map);
    if (m_dev) rq_device_close(m_dev);
}

// One device staging buffer for the gather's queries and results, kept
// between calls and grown on demand: a one-point GetNearestNPhotons then
// costs copies and a launch, not six allocations.
bool HipPhotonMap::ensureStage(size_t bytes) {
    if (bytes <= m_stageBytes) return true;
    if (m_stage) (void)hipFree(m_stage);
    m_stage = 0;
    m_stageBytes = 0;
    if (hipMalloc(&m_stage, bytes) != hipSuccess) {
        m_stage = 0;
        printf("HipPhotonMap: device staging allocation of %zu bytes failed\n", bytes);
        return false;
    }
    m_stageBytes = bytes;
    return true;
}

void HipPhotonMap::AddPhoton(Photon& p) {
    if (m_photons.size() < m_max) {
        m_photons.push_back(p);
        m_built = false;
    }
}

void HipPhotonMap::ScalePhotons(const float scaleFactor) {
    for (size_t i = 0; i < m_photons.size(); ++i) {
        Color c = m_photons[i].Power();
        c.R() *= scaleFactor;
        c.G() *= scaleFactor;
        c.B() *= scaleFactor;
        m_photons[i].Power(c);
    }
    m_built = false;
}

void HipPhotonMap::BalanceAndConstruct() {
    if (m_built) return;
    if (!m_dev && rq_device_open(m_ordinal, &m_dev) != RQ_OK) {
        printf("HipPhotonMap: %s\n", rq_last_error());
        return;
    }
    const size_t n = m_photons.size();
    std::vector<float> pos(3 * n), nrm(3 * n), pw(3 * n);
    for (size_t i = 0; i < n; ++i) {
        const Photon& p = m_photons[i];
        const Point3& x = p.Position();
        const Vector3& s = p.SurfNormal();
        const Color c = p.Power();
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] = x[k];
            nrm[3 * i + k] = s[k];
        }
        pw[3 * i] = c.R();
        pw[3 * i + 1] = c.G();
        pw[3 * i + 2] = c.B();
    }
    if (m_map) rq_photons_destroy(m_map);
    m_map = 0;
    if (rq_photons_create(m_dev, pos.data(), nrm.data(), pw.data(), (uint32_t)n, m_layout, &m_map) != RQ_OK) {
        printf("HipPhotonMap: %s\n", rq_last_error());
        return;
    }
    m_built = true;
}

// Every increment-th photon of ENCORE's balanced order gets the radiance
// estimate at its own position and normal as its power (the engine picks
// them and estimates on the device); the new map holds copies of those
// photons with that power, added in that order and balanced, as
// PhotonMap::PrecomputeIrradiance builds it.
HipPhotonMap* HipPhotonMap::PrecomputeIrradiance(int increment, const int N, const float maxDistance) {
    HipPhotonMap* out = new HipPhotonMap(m_max, m_ordinal, m_layout);
    if (!m_built) BalanceAndConstruct();
    if (!m_map || increment <= 0 || N <= 0) return out;
    const size_t n = m_photons.size();
    std::vector<float> pos(3 * n), nrm(3 * n);
    for (size_t i = 0; i < n; ++i) {
        const Point3& x = m_photons[i].Position();
        const Vector3& s = m_photons[i].SurfNormal();
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] = x[k];
            nrm[3 * i + k] = s[k];
        }
    }
    const size_t m = n ? (n - 1) / (size_t)increment + 1 : 0;
    std::vector<uint32_t> chosen(std::max<size_t>(1, m));
    std::vector<float> pw(3 * std::max<size_t>(1, m));
    rq_photons* made = 0;
    if (rq_photons_precompute_irradiance(m_map, pos.data(), nrm.data(), (uint32_t)n, (uint32_t)increment,
                                         (uint32_t)N, maxDistance, m_caustic ? 1u : 0u, m_layout, chosen.data(),
                                         pw.data(), &made) != RQ_OK) {
        printf("HipPhotonMap: %s\n", rq_last_error());
        return out;
    }
    rq_photons_destroy(made);  // the new map builds its own from its photons
    for (size_t j = 0; j < m; ++j) {
        Photon p = m_photons[chosen[j]];
        p.Power(Color(pw[3 * j], pw[3 * j + 1], pw[3 * j + 2]));
        out->AddPhoton(p);
    }
    out->BalanceAndConstruct();
    return out;
}

void HipPhotonMap::gather(int N, float maxDistance, const float* loc, const float* nrm, size_t nq, bool radiance,
                          std::vector<uint32_t>& idx, std::vector<float>& d2, std::vector<uint32_t>& cnt,
                          std::vector<float>& rgb) {
    idx.assign(nq * N, 0xFFFFFFFFu);
    d2.assign(nq * N, 0.0f);
    cnt.assign(nq, 0);
    rgb.assign(3 * nq, 0.0f);
    if (!m_built) BalanceAndConstruct();
    if (!m_map || nq == 0 || N <= 0) return;
    // Carve the staging buffer: queries, normals, indices, d2, counts, rgb.
    const size_t sizes[6] = {nq * 12, nq * 12, nq * (size_t)N * 4, nq * (size_t)N * 4, nq * 4, nq * 12};
    size_t off[6], total = 0;
    for (int i = 0; i < 6; ++i) {
        off[i] = total;
        total += (sizes[i] + 255) & ~(size_t)255;
    }
    bool ok = ensureStage(total);
    char* base = static_cast<char*>(m_stage);
    void *dq = base + off[0], *dn = base + off[1], *di = base + off[2], *dd = base + off[3], *dc = base + off[4],
         *dr = base + off[5];
    if (ok) {
        ok = hipMemcpy(dq, loc, nq * 12, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(dn, nrm, nq * 12, hipMemcpyHostToDevice) == hipSuccess;
    }
    rq_status s = RQ_OK;
    if (ok) {
        if (radiance)
            s = rq_radiance_estimate(m_map, (const float*)dq, (const float*)dn, (uint32_t)nq, (uint32_t)N, maxDistance,
                                     m_caustic ? 1u : 0u, (float*)dr, (uint32_t*)di, (float*)dd, (uint32_t*)dc, 0);
        else
            s = rq_knn(m_map, (const float*)dq, (const float*)dn, (uint32_t)nq, (uint32_t)N, maxDistance,
                       (uint32_t*)di, (float*)dd, (uint32_t*)dc, 0);
        if (s != RQ_OK) printf("HipPhotonMap: %s\n", rq_last_error());
    }
    if (ok && s == RQ_OK) {
        (void)hipMemcpy(idx.data(), di, nq * N * 4, hipMemcpyDeviceToHost);
        (void)hipMemcpy(d2.data(), dd, nq * N * 4, hipMemcpyDeviceToHost);
        (void)hipMemcpy(cnt.data(), dc, nq * 4, hipMemcpyDeviceToHost);
        if (radiance) (void)hipMemcpy(rgb.data(), dr, nq * 12, hipMemcpyDeviceToHost);
    }
}

void HipPhotonMap::GetNearestNPhotonsBatch(const int N, const float maxDistance, const std::vector<float>& locations,
                                           const std::vector<float>& normals,
                                           std::vector<std::vector<PhotonDistPair> >& out) {
    const size_t nq = locations.size() / 3;
    std::vector<uint32_t> idx, cnt;
    std::vector<float> d2, rgb;
    gather(N, maxDistance, locations.data(), normals.data(), nq, false, idx, d2, cnt, rgb);
    out.assign(nq, std::vector<PhotonDistPair>());
    for (size_t q = 0; q < nq; ++q) {
        std::vector<PhotonDistPair>& v = out[q];
        for (uint32_t j = 0; j < cnt[q]; ++j) v.push_back(PhotonDistPair(&m_photons[idx[q * N + j]], d2[q * N + j]));
        std::make_heap(v.begin(), v.end());  // PhotonMap.cpp:90
    }
}

std::vector<PhotonDistPair> HipPhotonMap::GetNearestNPhotons(const int N, const float maxDistance, Point3 location,
                                                             Vector3 normal) {
    std::vector<float> l(3), n(3);
    for (int k = 0; k < 3; ++k) {
        l[k] = location[k];
        n[k] = normal[k];
    }
    std::vector<std::vector<PhotonDistPair> > out;
    GetNearestNPhotonsBatch(N, maxDistance, l, n, out);
    return out.empty() ? std::vector<PhotonDistPair>() : out[0];
}

void HipPhotonMap::GetRadianceEstimateBatch(const int N, const float maxDistance, const std::vector<float>& locations,
                                            const std::vector<float>& normals, std::vector<Color>& out) {
    const size_t nq = locations.size() / 3;
    std::vector<uint32_t> idx, cnt;
    std::vector<float> d2, rgb;
    gather(N, maxDistance, locations.data(), normals.data(), nq, true, idx, d2, cnt, rgb);
    out.clear();
    for (size_t q = 0; q < nq; ++q) out.push_back(Color(rgb[3 * q], rgb[3 * q + 1], rgb[3 * q + 2]));
}

Color HipPhotonMap::GetRadianceEstimate(const int N, const float maxDistance, Point3 location, Vector3 normal) {
    std::vector<float> l(3), n(3);
    for (int k = 0; k < 3; ++k) {
        l[k] = location[k];
        n[k] = normal[k];
    }
    std::vector<Color> out;
    GetRadianceEstimateBatch(N, maxDistance, l, n, out);
    return out.empty() ? Color(0, 0, 0) : out[0];
}
