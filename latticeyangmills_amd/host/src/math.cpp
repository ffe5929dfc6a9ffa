// Host SU(3) algebra and the DeviceField registry.
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>

#include "Math/devicefield.h"
#include "Math/lattice.h"
#include "Math/su3.h"

SU3 &SU3::operator*=(const SU3 &o) noexcept
{
    SU3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double re = 0, im = 0;
            for (int k = 0; k < 3; k++) {
                const double ar = mat[2 * (3 * i + k)], ai = mat[2 * (3 * i + k) + 1];
                const double br = o.mat[2 * (3 * k + j)], bi = o.mat[2 * (3 * k + j) + 1];
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
            }
            r.mat[2 * (3 * i + j)] = re;
            r.mat[2 * (3 * i + j) + 1] = im;
        }
    return *this = r;
}

void SU3::setSU3Identity()
{
    mat.fill(0.0);
    mat[0] = mat[8] = mat[16] = 1.0;
}

void SU3::setSU3Zero() { mat.fill(0.0); }

void SU3::printSU3() const
{
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) printf("( %2.5g, %2.5g )\t", mat[2 * (3 * i + j)], mat[2 * (3 * i + j) + 1]);
        printf("\n");
    }
    printf("\n");
}

SU3 operator~(const SU3 &a)
{
    SU3 c;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            c.mat[2 * (3 * i + j)] = a.mat[2 * (3 * j + i)];
            c.mat[2 * (3 * i + j) + 1] = -a.mat[2 * (3 * j + i) + 1];
        }
    return c;
}

double getMultSumTrace(const SU3 &a, const SU3 &b)
{
    double s = 0;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++)
            s += a.mat[2 * (3 * i + k)] * b.mat[2 * (3 * k + i)] - a.mat[2 * (3 * i + k) + 1] * b.mat[2 * (3 * k + i) + 1];
    return s;
}

void setToZero(Lattice<SU3> &l)
{
    for (auto &m : l.lattice) m.setSU3Zero();
}

void setToIdentity(Lattice<SU3> &l)
{
    for (auto &m : l.lattice) m.setSU3Identity();
}

// ------------------------------------------------------------ DeviceField ---

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

DeviceField::Health DeviceField::health()
{
    toDevice();
    Health h{0.0, 0};
    check(lym_check_links(m_ctx, &h.maxDeviation, &h.nonFinite), "lym_check_links");
    return h;
}
