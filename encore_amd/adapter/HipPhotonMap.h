// HipPhotonMap.h -- drop-in for ENCORE's PhotonMap backed by the MI355X engine.
//
// Same interface and meaning as src/renderers/PhotonMap.h:28-104:
//   AddPhoton(Photon&)        store (host) until the map is full
//   BalanceAndConstruct()     build the balanced tree and upload it
//   GetNearestNPhotons(...)   vector<PhotonDistPair> in heap order (front =
//                             farthest), pointers into this map's photons,
//                             as PhotonMap.cpp:76-95 returns them
//   GetRadianceEstimate(...)  PhotonMap.cpp:121-164, caustic weighting when
//                             SetCausticsMap(true)
//   ScalePhotons(f)           scales stored power (before or after building)
//   PrecomputeIrradiance(...) PhotonMap.cpp:57-74: a new map (caller owns it)
// plus batched forms that take many gather points per call.
#pragma once

#include <cstdint>
#include <vector>

#include "Photon.h"
#include "PhotonMap.h"
#include "rq.h"

class HipPhotonMap {
public:
    explicit HipPhotonMap(size_t maxPhotonCount, int ordinal = 0,
                          rq_photon_layout layout = RQ_PHOTONS_SPATIAL);
    ~HipPhotonMap();

    void AddPhoton(Photon& p);
    void BalanceAndConstruct();
    std::vector<PhotonDistPair> GetNearestNPhotons(const int N, const float maxDistance, Point3 location,
                                                   Vector3 normal);
    Color GetRadianceEstimate(const int N, const float maxDistance, Point3 location, Vector3 normal);
    void ScalePhotons(const float scaleFactor);
    HipPhotonMap* PrecomputeIrradiance(int increment, const int N, const float maxDistance);
    bool Full() { return m_photons.size() == m_max; }
    int GetCountStored() { return (int)m_photons.size(); }
    size_t GetMaxPhotons() { return m_max; }
    void SetCausticsMap(bool c) { m_caustic = c; }

    // Batched: gather points (host arrays of 3*n floats); results per point.
    void GetNearestNPhotonsBatch(const int N, const float maxDistance, const std::vector<float>& locations,
                                 const std::vector<float>& normals, std::vector<std::vector<PhotonDistPair> >& out);
    void GetRadianceEstimateBatch(const int N, const float maxDistance, const std::vector<float>& locations,
                                  const std::vector<float>& normals, std::vector<Color>& out);

private:
    void gather(int N, float maxDistance, const float* loc, const float* nrm, size_t nq, bool radiance,
                std::vector<uint32_t>& idx, std::vector<float>& d2, std::vector<uint32_t>& cnt,
                std::vector<float>& rgb);
    bool ensureStage(size_t bytes);

    std::vector<Photon> m_photons;
    size_t m_max;
    bool m_caustic;
    bool m_built;
    int m_ordinal;
    rq_photon_layout m_layout;
    rq_device* m_dev;
    rq_photons* m_map;
    void* m_stage;        // device staging for gather(), kept between calls
    size_t m_stageBytes;
};
