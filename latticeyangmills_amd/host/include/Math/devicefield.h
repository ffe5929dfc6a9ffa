// DeviceField: the device copy of a host GluonField, one lym_ctx (include/lym.h).
// The GPU Action / Observable / App classes share it through a registry keyed
// by the GluonField's address, so the field crosses PCIe only when one side
// has changed: upload when the host copy was modified, download when a device
// operation (flow, Metropolis) has changed the links and the host reads them.
#pragma once
#include <array>
#include <memory>
#include <stdexcept>
#include <string>

#include "Math/field.h"
#include "lym.h"

class DeviceField {
public:
    // The DeviceField bound to `field` (created on first use with `precision`).
    static DeviceField &of(GluonField *field, int precision = LYM_PREC_FP64, int device = 0);
    // Drops the binding (frees the device memory).
    static void release(GluonField *field);

    ~DeviceField();
    DeviceField(const DeviceField &) = delete;
    DeviceField &operator=(const DeviceField &) = delete;

    lym_ctx *ctx() { return m_ctx; }
    GluonField *host() { return m_host; }
    int precision() const { return m_precision; }

    // The host links were written: they are now the current field and the
    // next device use uploads them.
    void hostModified()
    {
        m_hostNewer = true;
        m_deviceNewer = false;
    }
    // A device operation changed the links: the next host read downloads them.
    void deviceModified()
    {
        m_deviceNewer = true;
        m_hostNewer = false;
    }
    void toDevice();  // upload if the host copy is newer
    void toHost();    // download if the device copy is newer

    // Link health, reported and never fatal (lym_check_links): the largest
    // deviation of any link from SU(3) and the number of links holding a NaN
    // or an infinity.  Uploads first if the host copy is newer.
    struct Health {
        double maxDeviation;
        int64_t nonFinite;
    };
    Health health();

    // Picks the flow schedule for this lattice shape and precision once, by
    // timing the candidates on the current device field (lym_autotune); the
    // choice changes no result.  cachePath (may be empty) keeps the choice
    // across runs.  Returns the tuner's summary (empty after the first call).
    std::string autotune(const std::string &cachePath);

    // throws std::runtime_error with lym_last_error() on failure
    void check(int rc, const char *what);

private:
    DeviceField(GluonField *field, int precision, int device);
    GluonField *m_host;
    lym_ctx *m_ctx = nullptr;
    int m_precision;
    bool m_hostNewer = true, m_deviceNewer = false, m_tuned = false;
};
