// Cache of instantiated solver chunk plans, owned by a workspace.
#pragma once

#include <map>
#include <memory>
#include <string>

#include "workspace.hpp"

struct pm_workspace::PlanCache
{
    std::map<std::string, std::unique_ptr<pm::Plan>> plans;

    pm::Plan* find(const std::string& key)
    {
        auto it = plans.find(key);
        return it == plans.end() ? nullptr : it->second.get();
    }

    pm::Plan* insert(const std::string& key, std::unique_ptr<pm::Plan> plan)
    {
        if (plans.size() > 64) plans.clear();
        pm::Plan* raw = plan.get();
        plans[key] = std::move(plan);
        return raw;
    }
};
