// Minimal stand-in for lagrange::value_ptr (modules/core/include/lagrange/utils/value_ptr.h:133):
// an owning pointer; copying deep-copies the pointee.  Used only without the real Lagrange headers.
#pragma once

#include <memory>
#include <utility>

namespace lagrange {

template <typename T>
class value_ptr
{
public:
    value_ptr() = default;
    explicit value_ptr(T* p) : m_p(p) {}
    value_ptr(const value_ptr& o) : m_p(o.m_p ? new T(*o.m_p) : nullptr) {}
    value_ptr(value_ptr&&) noexcept = default;
    value_ptr& operator=(const value_ptr& o)
    {
        if (this != &o) m_p.reset(o.m_p ? new T(*o.m_p) : nullptr);
        return *this;
    }
    value_ptr& operator=(value_ptr&&) noexcept = default;
    T* operator->() const { return m_p.get(); }
    T& operator*() const { return *m_p; }
    T* get() const { return m_p.get(); }
    explicit operator bool() const { return static_cast<bool>(m_p); }

private:
    std::unique_ptr<T> m_p;
};

template <typename T, typename... Args>
value_ptr<T> make_value_ptr(Args&&... args)
{
    return value_ptr<T>(new T(std::forward<Args>(args)...));
}

} // namespace lagrange
