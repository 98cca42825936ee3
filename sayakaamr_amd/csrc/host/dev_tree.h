// dev_tree.h -- the host tree's tables to a device context (shared by MacSolver and FillPatch).
#pragma once

#include "amr_tree.h"

struct sayaka_dev_ctx;

namespace sayaka {
inline namespace amd {

// sayaka_dev_set_topology from the tree's nodes (AmrTreeNode, SayakaTree.h:52-82); throws on error
void upload_tree_topology(sayaka_dev_ctx* ctx, const AmrTree& tree);
// sayaka_dev_set_surroundings from AmrTree::surrBlocks (SayakaTree.h:336-526); throws on error
void upload_tree_surroundings(sayaka_dev_ctx* ctx, const AmrTree& tree);

}  // namespace amd
}  // namespace sayaka
