// The DeviceField registry: one device copy (lym_ctx) per host GluonField.
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

#include "Math/devicefield.h"

namespace {
std::map<GluonField *, std::unique_ptr<DeviceField>> &registry()
{
    static std::map<GluonField *, std::unique_ptr<DeviceField>> r;
    return r;
}
// Guards the registry only: a DeviceField itself, like its lym_ctx, is used
// from one thread at a time.
std::mutex &registry_lock()
{
    static std::mutex m;
    return m;
}
}  // namespace

DeviceField &DeviceField::of(GluonField *field, int precision, int device)
{
    if (!field) throw std::runtime_error("DeviceField: null GluonField");
    std::lock_guard<std::mutex> g(registry_lock());
    auto &r = registry();
    auto it = r.find(field);
    if (it == r.end()) it = r.emplace(field, std::unique_ptr<DeviceField>(new DeviceField(field, precision, device))).first;
    return *it->second;
}

void DeviceField::release(GluonField *field)
{
    std::lock_guard<std::mutex> g(registry_lock());
    registry().erase(field);
}

DeviceField::DeviceField(GluonField *field, int precision, int device) : m_host(field), m_precision(precision)
{
    int rc = lym_create(&m_ctx, field->m_size.data(), precision, device);
    if (rc != LYM_OK) throw std::runtime_error("lym_create failed with code " + std::to_string(rc));
}

DeviceField::~DeviceField() { lym_destroy(m_ctx); }

void DeviceField::check(int rc, const char *what)
{
    if (rc != LYM_OK) throw std::runtime_error(std::string(what) + ": " + lym_last_error(m_ctx));
}

void DeviceField::toDevice()
{
    if (!m_hostNewer) return;
    const double *p[4];
    for (int mu = 0; mu < 4; mu++) p[mu] = (*m_host)[mu].lattice.data()->mat.data();
    check(lym_upload_links(m_ctx, p), "lym_upload_links");
    m_hostNewer = false;
}

void DeviceField::toHost()
{
    if (!m_deviceNewer) return;
    double *p[4];
    for (int mu = 0; mu < 4; mu++) p[mu] = (*m_host)[mu].lattice.data()->mat.data();
    check(lym_download_links(m_ctx, p), "lym_download_links");
    m_deviceNewer = false;
}

std::string DeviceField::autotune(const std::string &cachePath)
{
    if (m_tuned) return "";
    toDevice();
    char summary[256] = {0};
    check(lym_autotune(m_ctx, cachePath.empty() ? nullptr : cachePath.c_str(), summary, int(sizeof summary)),
          "lym_autotune");
    m_tuned = true;
    return summary;
}

DeviceField::Health DeviceField::health()
{
    toDevice();
    Health h{0.0, 0};
    check(lym_check_links(m_ctx, &h.maxDeviation, &h.nonFinite), "lym_check_links");
    return h;
}
